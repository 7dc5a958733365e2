"""DiffusionSDE in the oracle (src/module/DiffusionSDE.cpp:48-254), on the reference's own random
stream (one MT19937 seeded once, consumed particle after particle, glibc libm): final positions,
directions, step sizes, lengths and times equal the reference's bit for bit in a uniform field, a
fine turbulent grid (the step search halves) and with advection (neutral particles drift)."""
import numpy as np
import pytest

import oracle_ctypes as O


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


ADVECTION_CASES = ["sde_adv_cartesian_shock", "sde_adv_spherical_shock_1d_cool", "sde_adv_spherical_shock_1d",
                   "sde_adv_oblique_shock", "sde_adv_spherical_advection_shock", "sde_adv_list"]


@pytest.mark.parametrize("name", ["sde_uniform", "sde_grid32", "sde_advection", "sde_advection_cspherical",
                                  "sde_advection_spherical"] + ADVECTION_CASES)
def test_oracle_sde_bit_exact_vs_reference(name):
    c = O.load_case(name)
    m = c["meta"]
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    O.sde_propagate(O.sde_field(c), st, int(c["nsteps"].max()), rng=O.SDE_MT, **O.sde_args(m))
    assert bits_equal(st.pos, c["final_pos"]) and bits_equal(st.dir, c["final_dir"])
    assert bits_equal(st.trajectory_length, c["final_len"]) and bits_equal(st.next_step, c["final_next"])
    assert bits_equal(st.time, c["final_time"])
    assert np.array_equal(st.active, c["final_active"])


def test_oracle_sde_threads_one_thread_is_the_reference_stream():
    """The CPU baseline's threaded form seeds thread t with seed + t: one thread is the reference."""
    c = O.load_case("sde_uniform")
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    n = O.sde_propagate_threads(O.sde_field(c), st, 10 ** 6, nthreads=1, **O.sde_args(c["meta"]))
    assert n == int(c["nsteps"].sum())
    assert bits_equal(st.pos, c["final_pos"]) and bits_equal(st.dir, c["final_dir"])


# ---------------------------------------------------------------- advection fields
ADVECTION_FIELDS = ["cartesian_shock", "spherical_shock_1d_cool", "spherical_shock_1d", "oblique_shock",
                    "spherical_advection_shock", "list"]


@pytest.mark.parametrize("name", ADVECTION_FIELDS)
def test_advection_fields_bit_exact_vs_reference(name):
    """getField of OneDimensionalCartesianShock, OneDimensionalSphericalShock (with and without upstream
    cooling), ObliqueAdvectionShock, SphericalAdvectionShock and an AdvectionFieldList (uniform + spherical
    shock + constant wind) at 2048 positions: the oracle with glibc equals the reference's samples bit for
    bit; the device's deterministic forms (tanh from the correctly rounded exp, e_phi without atan2/sin/cos)
    stay within 4e-16 of the largest velocity."""
    c = O.load_case("advection_fields")
    terms = c["meta"][name]
    ref = c["v_" + name]
    assert bits_equal(O.advection_eval(terms, c["pos"], O.LIBM_GLIBC), ref)
    det = O.advection_eval(terms, c["pos"], O.LIBM_CR)
    assert np.abs(det - ref).max() <= 4e-16 * np.abs(ref).max()
    assert np.mean(det == ref) > 0.7


def test_advection_shock_profiles_closed_forms():
    """The shock profiles' limits and symmetries: far upstream vUp, far downstream vUp / compressionRatio,
    the mean at the shock; the oblique shock's constant v_y; the radial shocks point along the position."""
    vUp, comp, L = 1e6, 4.0, 0.5 * O.kpc
    far = np.array([[-100 * L, 0, 0], [100 * L, 0, 0], [0.0, 0, 0]])
    for libm in (O.LIBM_GLIBC, O.LIBM_CR):
        v = O.advection_eval([(O.ADV_CARTESIAN_SHOCK, [comp, vUp, L])], far, libm)
        assert v[0, 0] == vUp and v[1, 0] == vUp / comp and v[2, 0] == (vUp + vUp / comp) * 0.5
        assert np.all(v[:, 1:] == 0)
        o = O.advection_eval([(O.ADV_OBLIQUE_SHOCK, [comp, vUp, 3e5, L])], far, libm)
        assert np.all(o[:, 1] == 3e5) and np.array_equal(o[:, 0], v[:, 0])
        rng = np.random.default_rng(5)
        p = rng.normal(size=(64, 3)) * 3 * O.kpc
        for kind, par in [(O.ADV_SPHERICAL_SHOCK_1D, [3 * O.kpc, vUp, comp, L, 1]),
                          (O.ADV_SPHERICAL_SHOCK_1D, [3 * O.kpc, vUp, comp, L, 0]),
                          (O.ADV_SPHERICAL_SHOCK, [0, 0, 0, 2 * O.kpc, 6e5, L, 3 * O.kpc, 0.0])]:
            w = O.advection_eval([(kind, par)], p, libm)
            cross = np.cross(w, p)
            assert np.abs(cross).max() <= 1e-12 * np.abs(w).max() * np.abs(p).max()
    # the list is the sum of its terms, from 0, in order
    c = O.load_case("advection_fields")
    terms = c["meta"]["list"]
    parts = [O.advection_eval([t], c["pos"]) for t in terms]
    assert bits_equal(O.advection_eval(terms, c["pos"]), (parts[0] + parts[1]) + parts[2])
