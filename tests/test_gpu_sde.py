"""DiffusionSDE on the GPU (crp_propagator_create_sde, src/module/DiffusionSDE.cpp:48-266).

Parity: the device's Philox stream against the oracle's Philox mode, bit for bit, on the inputs of
the reference fixtures (uniform field, a fine 32^3 grid where the step search halves, uniform
advection), in one call and split into several calls and launches.  The oracle's step logic is
pinned to the reference's own output on the reference's MT19937 stream (test_sde_oracle.py).
Statistics: the reference's tests (test/testDiffusionSDE.py) on the device's random numbers."""
import math

import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu

pc = 3.0856775814913673e16
kpc = 1e3 * pc
TeV, GeV, PeV = 1e12 * O.eV, 1e9 * O.eV, 1e15 * O.eV
c_light = 2.99792458e8


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def device_field(case):
    m = case["meta"]
    if "grid_spacing" in m:
        return capi.Field.grid3f(case["grid"], [0, 0, 0], m["grid_spacing"])
    return capi.Field.uniform(m["field"]["B"])


def sde_propagator(case, field, seed, launch=None):
    a = O.sde_args(case["meta"])
    p = capi.Propagator("SDE", field, a["tolerance"], a["min_step"], a["max_step"], max_length=a["max_len"],
                        epsilon=a["epsilon"], alpha=a["alpha"], scale=a["scale"], seed=seed,
                        advection=a.get("advection"))
    if launch:
        p.set_launch(**launch)
    return p


def compare(out, st, steps):
    assert bits_equal(out["pos"], st.pos)
    assert bits_equal(out["dir"], st.dir)
    assert bits_equal(out["next_step"], st.next_step)
    assert bits_equal(out["trajectory_length"], st.trajectory_length)
    assert bits_equal(out["time"], st.time)
    assert bits_equal(out["current_step"], st.current_step)
    assert np.array_equal(out["active"], st.active)
    assert np.array_equal(out["sde_step"], steps)


ADVECTION_CASES = ["sde_adv_cartesian_shock", "sde_adv_spherical_shock_1d_cool", "sde_adv_spherical_shock_1d",
                   "sde_adv_oblique_shock", "sde_adv_spherical_advection_shock", "sde_adv_list"]


@pytest.mark.parametrize("name", ["sde_uniform", "sde_grid32", "sde_advection", "sde_advection_cspherical",
                                  "sde_advection_spherical"] + ADVECTION_CASES)
@pytest.mark.parametrize("split", [None, (3, 5, -1), "launch"])
def test_sde_bit_exact_vs_oracle(name, split):
    c = O.load_case(name)
    m = c["meta"]
    seed = m["seed"]
    n = len(c["energy"])
    field = device_field(c)
    launch = dict(steps_per_launch=2, use_graph=True) if split == "launch" else None
    p = sde_propagator(c, field, seed, launch)
    b = capi.Batch(n)
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"])
    if split in (None, "launch"):
        p.propagate(b, max_steps=-1)
    else:
        for k in split:
            p.propagate(b, max_steps=k)
    out = b.download()
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    steps = np.zeros(n, dtype=np.uint32)
    O.sde_propagate(O.sde_field(c), st, 10 ** 6, rng=O.SDE_PHILOX, sde_step=steps, **O.sde_args(m))
    assert not st.active.any()
    compare(out, st, steps)


def test_sde_streams_and_counters_resume():
    """Stream numbers from the caller (64-bit, as serial numbers) and step counters carried from
    one upload to the next give the oracle's numbers."""
    c = O.load_case("sde_grid32")
    m = c["meta"]
    n = len(c["energy"])
    rng = np.random.default_rng(5)
    stream = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    start = rng.integers(0, 1000, n).astype(np.uint32)
    field = device_field(c)
    p = sde_propagator(c, field, 99)
    b = capi.Batch(n)
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"], sde_step=start, sde_stream=stream)
    p.propagate(b, max_steps=4)
    mid = b.download()
    b2 = capi.Batch(n)
    b2.upload(mid["pos"], mid["dir"], mid["energy"], mid["charge"], current_step=mid["current_step"],
              next_step=mid["next_step"], trajectory_length=mid["trajectory_length"], time=mid["time"],
              active=mid["active"], sde_step=mid["sde_step"], sde_stream=stream)
    p.propagate(b2, max_steps=7)
    out = b2.download()
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    steps = start.copy()
    a = O.sde_args(m)
    a["seed"] = 99
    O.sde_propagate(O.sde_field(c), st, 11, rng=O.SDE_PHILOX, sde_step=steps, sde_stream=stream, **a)
    compare(out, st, steps)


def test_sde_modulated_field_matches_oracle():
    """The same kernel over ModulatedMagneticFieldGrid (the kernel is templated on the field)."""
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    mc = O.load_case("ck_grid32_modulated")
    mm = mc["meta"]
    dfield = capi.Field.modulated_grid3f(g["grid"], gm["origin"], gm["spacing"], mc["mod"], mm["mod_origin"],
                                         mm["mod_spacing"], mreflective=mm["mod_reflective"])
    ofield = O.OracleField.modulated(g["grid"], gm["origin"], gm["spacing"], mc["mod"], mm["mod_origin"],
                                     mm["mod_spacing"], mreflective=mm["mod_reflective"])
    n = 96
    _, E, d = O.sample_particles(7, n, 1 * PeV, 100 * PeV)
    pos0 = np.asarray(gm["origin"]) + np.random.default_rng(7).uniform(0, 32 * gm["spacing"], (n, 3))
    q = np.full(n, O.eplus)
    q[::7] = 0
    # tolerance 1e-9: most steps halve their trial time several times
    kw = dict(tolerance=1e-9, min_step=1 * pc, max_step=10 * kpc, epsilon=0.2, alpha=1 / 3., scale=1.0,
              max_len=300 * kpc, seed=3)
    p = capi.Propagator("SDE", dfield, kw["tolerance"], kw["min_step"], kw["max_step"], max_length=kw["max_len"],
                        epsilon=kw["epsilon"], alpha=kw["alpha"], scale=kw["scale"], seed=kw["seed"])
    b = capi.Batch(n)
    b.upload(pos0, O.normalise(d), E, q)
    p.propagate(b, max_steps=-1)
    out = b.download()
    st = O.OracleState(pos0, d, E, q)
    steps = np.zeros(n, dtype=np.uint32)
    O.sde_propagate(ofield, st, 10 ** 6, rng=O.SDE_PHILOX, sde_step=steps, **kw)
    compare(out, st, steps)


def test_sde_refuses_mixed_precision():
    p = capi.Propagator("SDE", capi.Field.uniform([0, 0, 1e-13]), 1e-4, pc, kpc, max_length=kpc)
    with pytest.raises(capi.CrpError, match="FP64 only"):
        capi.check(capi.lib().crp_propagator_set_precision(p.handle, capi.PRECISION_MIXED))


# ---------------------------------------------------------------- the reference's tests
def anderson(x):
    """Anderson-Darling statistic for normality with estimated mean and sigma
    (test/testDiffusionSDE.py:31-139: Stephens' correction factor)."""
    y = np.sort(np.asarray(x, dtype=np.float64))
    N = len(y)
    w = (y - np.mean(y)) / np.std(y, ddof=1)
    cdf = np.array([0.5 * (1 + math.erf(v / math.sqrt(2))) for v in w])
    i = np.arange(1, N + 1)
    A2 = -N - np.sum((2 * i - 1.0) / N * (np.log(cdf) + np.log(1 - cdf)[::-1]))
    return A2 * (1 + 4. / N - 25. / N ** 2)


def run_transport(E, epsilon, adv, N=10 ** 4, seed=11):
    """N protons from the origin for 50 kpc in B = (0, 0, 1 nG): [DiffusionSDE, MaximumTrajectoryLength]."""
    L_max = 50 * kpc
    p = capi.Propagator("SDE", capi.Field.uniform([0, 0, 1e-13]), 1e-4, 1 * pc, 10 * kpc, max_length=L_max,
                        epsilon=epsilon, seed=seed, advection=adv)
    b = capi.Batch(N)
    b.upload(np.zeros((N, 3)), np.tile([-1.0, 0, 0], (N, 1)), np.full(N, E), np.full(N, O.eplus))
    st = p.propagate(b, max_steps=-1)
    out = b.download()
    assert not out["active"].any() and st.finished == N
    D = 1.0 * 6.1e24 * (E / (4 * GeV)) ** (1 / 3.)
    return out, D, L_max


@pytest.mark.parametrize("E", [10 * TeV, 1 * TeV, 10 * PeV])
def test_reference_parallel_diffusion_statistics(E):
    """DiffusionOneDirection.test_DiffusionEnergy*: z ~ normal with sigma sqrt(2 D L / c); no
    perpendicular motion at epsilon = 0 (tolerance of the reference: 4 standard errors)."""
    out, D, L = run_transport(E, 0.0, None)
    N = len(out["pos"])
    x, y, z = out["pos"].T
    assert np.all(x == 0) and np.all(y == 0)
    std_exp = np.sqrt(2 * D * L / c_light)
    assert anderson(z) < 2.274
    assert abs(np.mean(z) / 4. / (std_exp / np.sqrt(N))) < 1.
    assert abs((np.std(z) - std_exp) / 4. / (std_exp / np.sqrt(N) / np.sqrt(2.))) < 1.


def test_reference_full_transport_statistics():
    """DiffusionOneDirection.test_FullTransport: epsilon 0.1, advection 1e6 m/s along x."""
    eps, v = 0.1, 1e6
    out, D, L = run_transport(10 * TeV, eps, [v, 0, 0])
    N = len(out["pos"])
    x, y, z = out["pos"].T
    sx = np.sqrt(2 * eps * D * L / c_light)
    sz = np.sqrt(2 * D * L / c_light)
    for comp in (x, y, z):
        assert anderson(comp) < 2.274
    for comp, mean, std in ((x, v * L / c_light, sx), (y, 0.0, sx), (z, 0.0, sz)):
        assert abs((np.mean(comp) - mean) / 4. / (std / np.sqrt(N))) < 1.
        assert abs((np.std(comp) - std) / 4. / (std / np.sqrt(N) / np.sqrt(2.))) < 1.
