"""GPU parity of the FP64 path against the reference (golden fixtures) and the
oracle, called through the C ABI.

Bit-exactness bar:
  * fields and trajectories whose arithmetic has no transcendental call
    (Grid3f trilinear, uniform fields, fixed-step BP/CK in grids) are compared
    bit for bit with the reference's own output (tests/golden);
  * every case is compared bit for bit with the oracle run with the same
    correctly rounded cos/pow the kernels use (LIBM_CR); the oracle with
    glibc's libm is itself bit-exact against the golden files (test_oracle.py),
    so the only difference left between GPU and reference is glibc's rounding
    of pow/cos, quantified in test_oracle.py.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu

TRAJ_CASES = ["bp_fixed_grid32", "bp_adaptive_grid32", "ck_grid32_maxlen", "ck_fixed_grid32",
              "bp_fixed_planewave32", "ck_planewave32", "bp_fixed_uniform", "ck_uniform", "bp_fixed_grid64_c1",
              "bp_fixed_grid32_reflective", "ck_grid32_clip"]
# cases without a transcendental on the path: bit-exact against the reference itself
REF_EXACT = {"bp_fixed_grid32", "ck_fixed_grid32", "bp_fixed_uniform", "bp_fixed_grid64_c1",
             "bp_fixed_grid32_reflective"}


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def device_field(case):
    meta = case["meta"]
    f = meta.get("field")
    if f == "grid32":
        g = O.load_case("grid32")
        gm = g["meta"]["grid"]
        return capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"], reflective=meta.get("reflective", False),
                                 clip_volume=meta.get("clip", False))
    if isinstance(f, dict) and f["type"] == "planewave":
        t = O.planewave_tables(f)
        return capi.Field.planewave(t["k"], t["Ak"], t["xi"], t["kappa"], t["beta"])
    if isinstance(f, dict) and f["type"] == "uniform":
        return capi.Field.uniform(f["B"])
    gm = meta["grid"]
    return capi.Field.grid3f(O.noise_grid(gm["seed"], gm["n"], gm["amp"]), gm["origin"], gm["spacing"])


def make_prop(meta, field):
    a = O.propagator_args(meta)
    kind = "BP" if a["prop_type"] == O.PROP_BP else "CK"
    return capi.Propagator(kind, field, a["tolerance"], a["min_step"], a["max_step"],
                           max_length=meta.get("maxTrajectoryLength", 0.0))


# ------------------------------------------------------------------ fields
@pytest.mark.parametrize("mode,layout", [("B_periodic", "cells"), ("B_periodic", "points"),
                                         ("B_reflective", "points"), ("B_clip", "cells"), ("B_clip", "points")])
def test_trilinear_bit_exact(mode, layout):
    c = O.load_case("trilinear16")
    m = c["meta"]
    f = capi.Field.grid3f(c["grid"], m["origin"], m["spacing"], reflective=(mode == "B_reflective"),
                          clip_volume=(mode == "B_clip"))
    f.set_layout(layout)
    assert f.layout() == layout
    B = f.eval(c["pos"])
    assert bits_equal(B, c[mode])


def test_grid_layout_defaults():
    c = O.load_case("trilinear16")
    m = c["meta"]
    assert capi.Field.grid3f(c["grid"], m["origin"], m["spacing"]).layout() == "cells"
    fr = capi.Field.grid3f(c["grid"], m["origin"], m["spacing"], reflective=True)
    assert fr.layout() == "points"
    with pytest.raises(capi.CrpError):
        fr.set_layout("cells")


def test_nonuniform_grid_shape_cells():
    """Non-cubic grid with per-axis spacing: both layouts equal the oracle bit for bit."""
    rng = np.random.default_rng(4)
    g = (rng.normal(size=(5, 7, 3, 3)) * 1e-13).astype(np.float32)
    origin, sp = [-1e19, 2e19, 0.0], [3e19, 1e19, 4e19]
    pos = rng.uniform(-5e20, 5e20, size=(3000, 3))
    ref = O.OracleField.grid(g, origin, sp).eval(pos)
    for layout in ("cells", "points"):
        f = capi.Field.grid3f(g, origin, sp).set_layout(layout)
        assert bits_equal(f.eval(pos), ref), layout


@pytest.mark.parametrize("name", ["planewave_c3", "planewave_general10", "planewave_simple2"])
def test_planewave_field(name):
    c = O.load_case(name)
    t = O.planewave_tables(c["meta"])
    f = capi.Field.planewave(t["k"], t["Ak"], t["xi"], t["kappa"], t["beta"])
    B = f.eval(c["pos"])
    ref_cr = O.OracleField.planewave(t).eval(c["pos"], libm=O.LIBM_CR)
    assert bits_equal(B, ref_cr)
    # against glibc's cos (the reference): only last-bit differences of single modes
    brms = c["meta"]["Brms"]
    assert np.max(np.abs(B - c["B"])) <= 1e-13 * brms


def test_uniform_and_null_field():
    f = capi.Field.uniform([1e-9, -2e-10, 3e-11])
    pos = np.random.default_rng(0).normal(size=(100, 3))
    assert np.all(f.eval(pos) == np.array([1e-9, -2e-10, 3e-11]))


# ------------------------------------------------------------------ trajectories
@pytest.mark.parametrize("name", TRAJ_CASES + ["bp_fixed_grid32:points", "ck_grid32_maxlen:points",
                                  "ck_grid32_clip:points"])
def test_trajectory_parity(name):
    name, _, layout = name.partition(":")
    c = O.load_case(name)
    m = c["meta"]
    field = device_field(c)
    if layout:
        field.set_layout(layout)
    prop = make_prop(m, field)
    n = len(c["energy"])
    b = capi.Batch(n)
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"])
    # oracle with the kernels' libm, advanced in the same chunks
    of = O.field_for_case(c)
    ost = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    pa = O.propagator_args(m)
    done = 0
    for si, s in enumerate(c["snap_steps"]):
        prop.propagate(b, max_steps=int(s) - done)
        O.propagate(of, steps=int(s) - done, max_len=m.get("maxTrajectoryLength", 0.0), libm=O.LIBM_CR,
                    state=ost, **pa)
        done = int(s)
        g = b.download()
        for key, ref in (("pos", ost.pos), ("dir", ost.dir), ("next_step", ost.next_step),
                         ("current_step", ost.current_step), ("trajectory_length", ost.trajectory_length),
                         ("time", ost.time)):
            assert bits_equal(g[key], ref), (name, s, key)
        assert np.array_equal(g["active"], ost.active)
        if name in REF_EXACT:
            assert bits_equal(g["pos"], c["pos"][si]), (name, s)
            assert bits_equal(g["dir"], c["dir"][si]), (name, s)
            assert bits_equal(g["trajectory_length"], c["trajectory_length"][si])
            assert bits_equal(g["time"], c["time"][si])
            assert bits_equal(g["next_step"], c["next_step"][si])
    assert bits_equal(g["prev_pos"], ost.prev_pos)
    assert bits_equal(g["prev_dir"], ost.prev_dir)
    if name in REF_EXACT:
        assert bits_equal(g["prev_pos"], c["final_previous_pos"])


def _c2_small(n, seed=7):
    """C2-shaped workload at a size the oracle finishes in seconds."""
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    sp, E, d = O.sample_particles(seed, n, 1 * O.EeV, 300 * O.EeV, weights=[0.5, 0.2, 0.15, 0.1, 0.05])
    Z = np.array([1, 2, 7, 14, 26])[sp]
    q = Z * O.eplus
    L = gm["n"] * gm["spacing"]
    pos = np.tile([L / 2, L / 2, L / 2], (n, 1))
    return g["grid"], gm, pos, d, E, q


@pytest.mark.parametrize("kind,tol,mins,maxs", [("CK", 1e-4, 1 * O.kpc, 1 * O.Mpc),
                                                ("BP", 1e-3, 1 * O.kpc, 1 * O.Mpc)])
def test_adaptive_to_max_length_matches_oracle(kind, tol, mins, maxs):
    grid, gm, pos, d, E, q = _c2_small(4096)
    field = capi.Field.grid3f(grid, gm["origin"], gm["spacing"])
    prop = capi.Propagator(kind, field, tol, mins, maxs, max_length=3 * O.Mpc)
    b = capi.Batch(len(E))
    b.upload(pos, O.normalise(d), E, q)
    st = prop.propagate(b, max_steps=-1)
    g = b.download()
    of = O.OracleField.grid(grid, gm["origin"], gm["spacing"])
    ost = O.OracleState(pos, d, E, q)
    nacc, _ = O.propagate(of, O.PROP_CK if kind == "CK" else O.PROP_BP, ost, 10 ** 9, tolerance=tol, min_step=mins,
                          max_step=maxs, max_len=3 * O.Mpc, libm=O.LIBM_CR)
    assert st.accepted_steps == nacc
    assert st.finished == len(E)
    assert np.all(g["active"] == 0)
    for key in ("pos", "dir", "trajectory_length", "time", "next_step", "current_step"):
        assert bits_equal(g[key], getattr(ost, key)), key
    assert np.all(g["trajectory_length"] >= 3 * O.Mpc)


def test_planewave_bp_matches_oracle():
    c = O.load_case("planewave_c3")
    t = O.planewave_tables(c["meta"])
    t64 = {k: v[:64] for k, v in t.items()}  # first 64 modes of the C3 table
    field = capi.Field.planewave(t64["k"], t64["Ak"], t64["xi"], t64["kappa"], t64["beta"])
    n = 1024
    _, E, d = O.sample_particles(11, n, 1e15 * O.eV, 1e18 * O.eV)
    q = np.full(n, O.eplus)
    pos = np.zeros((n, 3))
    prop = capi.Propagator("BP", field, 0.42, 0.05 * O.pc, 0.05 * O.pc)
    b = capi.Batch(n)
    b.upload(pos, O.normalise(d), E, q)
    prop.propagate(b, max_steps=200)
    g = b.download()
    ost = O.OracleState(pos, d, E, q)
    O.propagate(O.OracleField.planewave(t64), O.PROP_BP, ost, 200, min_step=0.05 * O.pc, max_step=0.05 * O.pc,
                libm=O.LIBM_CR)
    assert bits_equal(g["pos"], ost.pos)
    assert bits_equal(g["dir"], ost.dir)


# ------------------------------------------------------------------ edge cases
def test_empty_batch_and_zero_steps():
    field = capi.Field.uniform([0, 0, 1e-13])
    prop = capi.Propagator("BP", field, 0.42, O.kpc, O.kpc)
    b = capi.Batch(16)
    b.upload(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    st = prop.propagate(b, max_steps=10)
    assert st.accepted_steps == 0
    b.upload(np.zeros((4, 3)), np.tile([0, 1.0, 0], (4, 1)), np.full(4, O.EeV), np.full(4, O.eplus))
    st = prop.propagate(b, max_steps=0)
    assert st.accepted_steps == 0
    assert np.all(b.download()["pos"] == 0)


def test_inactive_particles_untouched_and_ragged_budget():
    field = capi.Field.uniform([0, 0, 1e-13])
    prop = capi.Propagator("CK", field, 1e-4, 0.1 * O.kpc, 1 * O.Mpc, max_length=10 * O.kpc)
    n = 1000
    rng = np.random.default_rng(3)
    d = O.normalise(rng.normal(size=(n, 3)))
    E = 10 ** rng.uniform(17, 20, n) * O.eV
    act = (rng.uniform(size=n) > 0.3).astype(np.int32)
    b = capi.Batch(n)
    b.upload(np.zeros((n, 3)), d, E, np.full(n, O.eplus), active=act)
    st = prop.propagate(b, max_steps=-1)
    g = b.download()
    assert np.all(g["pos"][act == 0] == 0)
    assert np.all(g["trajectory_length"][act == 0] == 0)
    assert np.all(g["trajectory_length"][act == 1] >= 10 * O.kpc)
    assert st.finished == act.sum()


def test_neutral_rectilinear():
    # test/testPropagation.cpp:252-270 (CK) and :529-548 (BP)
    field = capi.Field.uniform([0, 0, 1 * O.nG])
    for kind, mx, nxt in (("CK", 42 * O.Mpc, 42 * O.Mpc), ("BP", 1 * O.kpc, 1 * O.kpc)):
        prop = capi.Propagator(kind, field, 1e-4 if kind == "CK" else 0.42, 1 * O.kpc, mx)
        b = capi.Batch(1)
        b.upload([[0, 0, 0]], [[0, 1, 0]], [100 * O.EeV], [0.0])
        prop.propagate(b, max_steps=1)
        g = b.download()
        assert g["current_step"][0] == 1 * O.kpc
        assert g["next_step"][0] == nxt
        assert np.array_equal(g["pos"][0], [0, 1 * O.kpc, 0])
        assert np.array_equal(g["dir"][0], [0, 1, 0])


def test_unbounded_run_rejected():
    field = capi.Field.uniform([0, 0, 1e-13])
    prop = capi.Propagator("BP", field, 0.42, O.kpc, O.kpc)
    b = capi.Batch(4)
    b.upload(np.zeros((4, 3)), np.tile([1.0, 0, 0], (4, 1)), np.full(4, O.EeV), np.full(4, O.eplus))
    with pytest.raises(capi.CrpError):
        prop.propagate(b, max_steps=-1)
