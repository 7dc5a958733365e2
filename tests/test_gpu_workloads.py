"""The benchmark workloads themselves under a check, at their real grids.

* ck_grid256 (adaptive Cash-Karp to 100 Mpc in the 256^3 Kolmogorov grid of
  seed 7, nuclei Z in {1,2,7,14,26}): the first 2048 particles the benchmark
  draws, FP64 GPU bit for bit against the oracle with the correctly rounded
  pow the kernel uses, every particle run to the length limit.
* bp_grid512 / ck_grid512 (fixed 20 kpc steps in the 512^3 grid of seed 3,
  10 kpc spacing, 2^27 cells): FP64 GPU bit for bit against the oracle with
  glibc (no transcendental is on a fixed-step grid path, so this is the
  reference's arithmetic), particles spread over and beyond the whole box so
  every part of the cell table and the periodic wrap are read.
* the mixed path on the 512^3 grid, FP32 cells and FP16 bricks, Boris and
  Cash-Karp: per-particle position-error percentiles at 10 and 100 steps
  (bounded), and the running diffusion coefficients D_perp, D_par after 1000
  steps against the same particles' FP64 run, as a paired difference within
  4 standard errors (a bias of a few percent in either coefficient fails it).
"""
import numpy as np
import pytest

import bench
import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rel_err(a, b):
    return np.max(np.abs(a - b), axis=1) / np.max(np.abs(b), axis=1)


def pct(e):
    return "median %.2e p90 %.2e p99 %.2e max %.2e" % (np.median(e), np.percentile(e, 90), np.percentile(e, 99),
                                                      e.max())


def test_ck_grid256_workload_matches_oracle():
    w = dict(bench.WORKLOADS["ck_grid256"])
    n = 2048
    grid = bench.make_field_data(w)
    assert grid.shape == (256, 256, 256, 3) and grid.dtype == np.float32
    pos, d, E, q = bench.make_particles(w, n)
    sp = [w["spacing"]] * 3
    print("ck_grid256: 2048 particles, CK 1e-4 / 1 kpc / 1 Mpc to 100 Mpc, FP64", flush=True)
    field = capi.Field.grid3f(grid, [0, 0, 0], sp)
    prop = capi.Propagator("CK", field, w["tolerance"], w["min_step"], w["max_step"], max_length=w["max_length"])
    b = capi.Batch(n)
    b.upload(pos, d, E, q)
    st = prop.propagate(b, max_steps=-1)
    g = b.download()
    ost = O.OracleState(pos, d, E, q, normalised=True)
    nacc, trials = O.propagate(O.OracleField.grid(grid, [0, 0, 0], sp), O.PROP_CK, ost, 10 ** 9,
                               tolerance=w["tolerance"], min_step=w["min_step"], max_step=w["max_step"],
                               max_len=w["max_length"], libm=O.LIBM_CR)
    print("accepted steps GPU %d oracle %d, trials %d" % (st.accepted_steps, nacc, trials))
    assert st.accepted_steps == nacc
    assert st.finished == n and not g["active"].any()
    for key in ("pos", "dir", "trajectory_length", "time", "next_step", "current_step", "prev_pos", "prev_dir"):
        assert bits_equal(g[key], getattr(ost, key)), key
    assert np.all(g["trajectory_length"] >= w["max_length"] * (1 - 1e-15))


@pytest.fixture(scope="module")
def grid512():
    w = dict(bench.WORKLOADS["bp_grid512"])
    print("generating the 512^3 grid (seed 3)", flush=True)
    g = bench.make_field_data(w)
    assert g.shape == (512, 512, 512, 3)
    return w, g


@pytest.mark.parametrize("kind", ["BP", "CK"])
@pytest.mark.parametrize("layout", ["cells", "points"])
def test_grid512_fixed_step_fp64_matches_oracle(grid512, kind, layout):
    w, grid = grid512
    n, steps = 1024, 200
    _, E, d = bench.draw_particles(w["seed"], n, w["emin"], w["emax"])
    L = 512 * w["spacing"]
    pos = np.random.default_rng(5).uniform(-L, 2 * L, size=(n, 3))
    pos[:64] = L - 1e-3 * w["spacing"]  # right at the upper edge: the wrap to cell 0
    q = np.full(n, bench.eplus)
    sp = [w["spacing"]] * 3
    print("grid512 %s %s: %d particles x %d steps" % (kind, layout, n, steps), flush=True)
    field = capi.Field.grid3f(grid, [0, 0, 0], sp).set_layout(layout)
    tol = 0.42 if kind == "BP" else 1e-4
    h = w["min_step"]
    prop = capi.Propagator(kind, field, tol, h, h)
    b = capi.Batch(n)
    b.upload(pos, O.normalise(d), E, q)
    st = prop.propagate(b, max_steps=steps)
    g = b.download()
    ost = O.OracleState(pos, d, E, q)
    O.propagate(O.OracleField.grid(grid, [0, 0, 0], sp), O.PROP_BP if kind == "BP" else O.PROP_CK, ost, steps,
                tolerance=tol, min_step=h, max_step=h, libm=O.LIBM_GLIBC)
    assert st.accepted_steps == n * steps
    for key in ("pos", "dir", "trajectory_length", "prev_pos", "prev_dir"):
        assert bits_equal(g[key], getattr(ost, key)), key
    # the particles crossed the box many times over: every octant of the table was read
    cells = np.floor(np.mod(g["pos"], L) / w["spacing"]).astype(int)
    assert len(np.unique((cells >= 256).astype(int) @ [4, 2, 1])) == 8


def paired_D(disp_a, disp_b, L):
    """Running diffusion coefficients of run a minus run b on the same particles, with the standard error
    of the paired difference: D_perp = <dx^2 + dy^2>/(4L), D_par = <dz^2>/(2L)."""
    out = []
    for sel, den in (((0, 1), 4 * L), ((2,), 2 * L)):
        a = np.sum(disp_a[:, sel] ** 2, axis=1) / den
        b = np.sum(disp_b[:, sel] ** 2, axis=1) / den
        diff = a - b
        out.append((a.mean(), b.mean(), diff.mean(), diff.std(ddof=1) / np.sqrt(len(diff))))
    return out


# per-particle relative position error bounds (median, p90) at 10 and 100 steps, about ten times the
# measured values (BP / CK, median and p90: FP32 cells 10 steps 2.3e-9 4.3e-9 / 3.1e-9 6.1e-9, 100 steps
# 3.5e-8 8.4e-8 / 6.4e-8 1.6e-7; FP16 bricks 10 steps 2.2e-8 1.6e-7 / 1.7e-8 1.2e-7, 100 steps
# 3.7e-7 2.8e-6 / 3.2e-7 2.3e-6)
MIXED_BOUNDS = {("fp32", 10): (3e-8, 6e-8), ("fp32", 100): (6e-7, 2e-6),
                ("fp16", 10): (3e-7, 2e-6), ("fp16", 100): (4e-6, 3e-5)}


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("kind", ["BP", "CK"])
def test_grid512_mixed_vs_fp64(grid512, kind, storage):
    w, grid = grid512
    n = 1 << 14
    w["n"] = n
    pos, d, E, q = bench.make_particles(w, n)
    sp = [w["spacing"]] * 3
    h = w["min_step"]
    tol = 0.42 if kind == "BP" else 1e-4
    field = capi.Field.grid3f(grid, [0, 0, 0], sp)
    if storage == "fp16":
        field.set_storage("fp16")
    print("grid512 mixed %s %s storage: %d particles" % (kind, storage, n), flush=True)
    runs = {}
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        prop = capi.Propagator(kind, field, tol, h, h, precision=prec)
        b = capi.Batch(n)
        b.upload(pos, d, E, q)
        done, out = 0, {}
        for s in (10, 100, 1000):
            prop.propagate(b, max_steps=s - done)
            done = s
            out[s] = b.download()
        runs[prec] = out
    ref, mix = runs[capi.PRECISION_FP64], runs[capi.PRECISION_MIXED]
    for s in (10, 100):
        e = rel_err(mix[s]["pos"], ref[s]["pos"])
        print("%s %s %d steps: relative position error %s" % (kind, storage, s, pct(e)))
        med, p90 = MIXED_BOUNDS[(storage, s)]
        assert np.median(e) < med and np.percentile(e, 90) < p90
    e = rel_err(mix[1000]["pos"], ref[1000]["pos"])
    print("%s %s 1000 steps: relative position error %s" % (kind, storage, pct(e)))
    L = 1000 * h
    for name, (a, b_, diff, se) in zip(("perp", "par"), paired_D(mix[1000]["pos"] - pos, ref[1000]["pos"] - pos, L)):
        print("%s %s D_%s mixed %.5e fp64 %.5e: difference %.2e = %.2f standard errors (%.2e relative)"
              % (kind, storage, name, a, b_, diff, diff / se, diff / b_))
        assert abs(diff) <= 4 * se
    assert np.allclose(np.linalg.norm(mix[1000]["dir"], axis=1), 1, atol=1e-6)
