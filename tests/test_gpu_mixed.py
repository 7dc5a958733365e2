"""Mixed precision (FP32 field and rotation, FP64 positions and step control):
error against the FP64 path, stated as bounds on the per-particle relative
final-position error."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.max(np.abs(a - b), axis=1) / np.max(np.abs(b), axis=1)


def run(field, kind, tol, mn, mx, pos, d, E, q, steps, precision, max_len=0.0):
    p = capi.Propagator(kind, field, tol, mn, mx, max_length=max_len, precision=precision)
    b = capi.Batch(len(E))
    b.upload(pos, d, E, q)
    st = p.propagate(b, max_steps=steps)
    return b.download(), st


def grid_setup(n=2048, seed=42):
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    _, E, d = O.sample_particles(seed, n, 1 * O.EeV, 100 * O.EeV)
    L = gm["n"] * gm["spacing"]
    return g["grid"], gm, np.tile([L / 2] * 3, (n, 1)), O.normalise(d), E, np.full(n, O.eplus)


@pytest.mark.parametrize("layout", ["cells", "points"])
def test_mixed_bp_fixed_grid(layout):
    grid, gm, pos, d, E, q = grid_setup()
    f = capi.Field.grid3f(grid, gm["origin"], gm["spacing"]).set_layout(layout)
    ref, _ = run(f, "BP", 0.42, 0.1 * O.Mpc, 0.1 * O.Mpc, pos, d, E, q, 1000, capi.PRECISION_FP64)
    mix, _ = run(f, "BP", 0.42, 0.1 * O.Mpc, 0.1 * O.Mpc, pos, d, E, q, 1000, capi.PRECISION_MIXED)
    e = rel_err(mix["pos"], ref["pos"])
    print("BP fixed 1000 steps: rel. position error median %.2e p90 %.2e max %.2e" %
          (np.median(e), np.percentile(e, 90), e.max()))
    assert np.median(e) < 1e-5
    assert np.allclose(np.linalg.norm(mix["dir"], axis=1), 1, atol=1e-6)
    assert np.array_equal(mix["trajectory_length"], ref["trajectory_length"])


def test_mixed_ck_grid_short_horizon():
    grid, gm, pos, d, E, q = grid_setup()
    f = capi.Field.grid3f(grid, gm["origin"], gm["spacing"])
    ref, sr = run(f, "CK", 1e-4, 1 * O.kpc, 1 * O.Mpc, pos, d, E, q, 100, capi.PRECISION_FP64)
    mix, sm = run(f, "CK", 1e-4, 1 * O.kpc, 1 * O.Mpc, pos, d, E, q, 100, capi.PRECISION_MIXED)
    e = rel_err(mix["pos"], ref["pos"])
    print("CK 100 steps: rel. position error median %.2e p90 %.2e max %.2e; trials %d vs %d" %
          (np.median(e), np.percentile(e, 90), e.max(), sm.rejected_trials, sr.rejected_trials))
    assert np.median(e) < 1e-4
    assert abs(sm.rejected_trials - sr.rejected_trials) <= 0.02 * max(sr.rejected_trials, 1) + 10


def test_mixed_uniform_gyration():
    f = capi.Field.uniform([0, 0, 1 * O.nG])
    step = 10 * O.Mpc
    out, _ = run(f, "BP", 0.42, step, step, [[0, 0, 0]], O.normalise([[1, 1, 1]]), [100 * O.EeV], [O.eplus], 10,
                 capi.PRECISION_MIXED)
    d, p = out["dir"][0], out["pos"][0]
    assert d[0] ** 2 + d[1] ** 2 == pytest.approx(2 / 3., rel=1e-6)
    assert p[2] ** 2 == pytest.approx(100 * step * step / 3., rel=1e-6)


def test_mixed_planewave_field_and_bp():
    c = O.load_case("planewave_c3")
    t = O.planewave_tables(c["meta"])
    f = capi.Field.planewave(t["k"], t["Ak"], t["xi"], t["kappa"], t["beta"])
    n = 1024
    _, E, d = O.sample_particles(11, n, 1e15 * O.eV, 1e18 * O.eV)
    q = np.full(n, O.eplus)
    pos = np.zeros((n, 3))
    ref, _ = run(f, "BP", 0.42, 0.05 * O.pc, 0.05 * O.pc, pos, O.normalise(d), E, q, 1000, capi.PRECISION_FP64)
    mix, _ = run(f, "BP", 0.42, 0.05 * O.pc, 0.05 * O.pc, pos, O.normalise(d), E, q, 1000, capi.PRECISION_MIXED)
    e = rel_err(mix["pos"], ref["pos"])
    print("plane waves BP 1000 steps: rel. position error median %.2e p90 %.2e max %.2e" %
          (np.median(e), np.percentile(e, 90), e.max()))
    assert np.median(e) < 1e-4


def test_mixed_ck_to_max_length_completes():
    grid, gm, pos, d, E, q = grid_setup(4096, 7)
    f = capi.Field.grid3f(grid, gm["origin"], gm["spacing"])
    mix, st = run(f, "CK", 1e-4, 1 * O.kpc, 1 * O.Mpc, pos, d, E, q, -1, capi.PRECISION_MIXED, max_len=3 * O.Mpc)
    assert np.all(mix["active"] == 0)
    assert st.finished == len(E)
    assert np.all(mix["trajectory_length"] >= 3 * O.Mpc)
