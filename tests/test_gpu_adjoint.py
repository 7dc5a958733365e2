"""Adjoint of the Boris push through a periodic grid (crp_adjoint_*), behind
torch.autograd: checked against the oracle's FP64 CPU reference gradient (an
independent implementation that tapes every state instead of checkpointing,
itself pinned by finite differences) and against central finite differences
of the GPU forward pass."""
import numpy as np
import pytest

import oracle_ctypes as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def setup(n=64, N=16, nsteps=300, K=32, seed=3):
    from crpropa3_amd.turbulence import simple_grid_turbulence
    sp = 20 * O.kpc
    grid = simple_grid_turbulence(N, sp, 1 * O.nG, 2 * sp, N * sp, seed).astype(np.float32)
    rng = np.random.default_rng(seed)
    _, E, d = O.sample_particles(seed + 1, n, 1 * O.EeV, 100 * O.EeV)
    u0 = O.normalise(d)
    x0 = rng.uniform(0, N * sp, size=(n, 3))
    target = O.normalise(rng.normal(size=(n, 3)))
    wx = rng.normal(size=(n, 3)) / (N * sp)
    h = 10 * O.kpc
    kq = O.eplus * O.c_light * h / (2 * E)
    return dict(grid=grid, sp=sp, x0=x0, u0=u0, E=E, kq=kq, target=target, wx=wx, h=h, nsteps=nsteps, K=K, N=N)


def oracle_grad(s, scale=1.0):
    o, sp = np.zeros(3), np.full(3, s["sp"])
    xo, uo = O.adjoint_forward(s["grid"], o, sp, scale, s["h"], s["nsteps"], s["x0"], s["u0"], s["kq"])
    gg, gx0, gu0, gs = O.adjoint_backward(s["grid"], o, sp, scale, s["h"], s["nsteps"], s["x0"], s["u0"], s["kq"],
                                          s["wx"], 2 * (uo - s["target"]))
    loss = np.sum((uo - s["target"]) ** 2) + np.sum(s["wx"] * xo)
    return loss, xo, uo, gg, gx0, gu0, gs


def gpu_run(s, scale=1.0, grad=True, grid=None, u0=None):
    from crpropa3_amd.adjoint import boris_grid_propagate
    dev = torch.device("cuda:0")
    g = torch.tensor(s["grid"] if grid is None else grid, dtype=torch.float32, device=dev, requires_grad=grad)
    x0 = torch.tensor(s["x0"], dtype=torch.float64, device=dev, requires_grad=grad)
    u0 = torch.tensor(s["u0"] if u0 is None else u0, dtype=torch.float64, device=dev, requires_grad=grad)
    sc = torch.tensor(scale, dtype=torch.float64, device=dev, requires_grad=grad)
    tgt = torch.tensor(s["target"], dtype=torch.float64, device=dev)
    wx = torch.tensor(s["wx"], dtype=torch.float64, device=dev)
    with torch.set_grad_enabled(grad):
        xo, uo = boris_grid_propagate(g, x0, u0, torch.tensor(s["E"], device=dev), O.eplus, [0, 0, 0],
                                      [s["sp"]] * 3, s["h"], s["nsteps"], s["K"], scale=sc)
        loss = ((uo - tgt) ** 2).sum() + (wx * xo).sum()
    if not grad:
        return loss.item()
    loss.backward()
    torch.cuda.synchronize()
    return (loss.item(), xo.detach().cpu().numpy(), uo.detach().cpu().numpy(), g.grad.double().cpu().numpy(),
            x0.grad.cpu().numpy(), u0.grad.cpu().numpy(), sc.grad.item())


def test_adjoint_forward_and_gradients_match_cpu_reference():
    s = setup()
    ref = oracle_grad(s)
    got = gpu_run(s)
    print("loss %.6e, forward max |dx| %.2e m, grid-gradient max rel %.2e" % (
        got[0], np.abs(got[1] - ref[1]).max(), np.abs(got[3] - ref[3]).max() / np.abs(ref[3]).max()))
    # the GPU divides by multiplying with reciprocals: agreement to rounding, not bit-exact
    assert got[0] == pytest.approx(ref[0], rel=1e-10)
    assert np.allclose(got[1], ref[1], rtol=1e-10, atol=0)
    assert np.allclose(got[2], ref[2], rtol=0, atol=1e-10)
    # the grid gradient comes back in the grid's dtype (float32)
    assert np.max(np.abs(got[3] - ref[3])) < 1e-6 * np.abs(ref[3]).max()
    assert np.allclose(got[4], ref[4], rtol=1e-7, atol=1e-8 * np.abs(ref[4]).max())
    assert np.allclose(got[5], ref[5], rtol=1e-7, atol=1e-8 * np.abs(ref[5]).max())
    assert got[6] == pytest.approx(ref[6], rel=1e-8)


def test_adjoint_partial_segment_and_no_position_adjoint():
    """nsteps not a multiple of K; loss on directions only (no adjoint of x_final)."""
    from crpropa3_amd.adjoint import boris_grid_propagate
    s = setup(n=100, nsteps=77, K=10)
    o, sp = np.zeros(3), np.full(3, s["sp"])
    xo, uo = O.adjoint_forward(s["grid"], o, sp, 1.0, s["h"], s["nsteps"], s["x0"], s["u0"], s["kq"])
    gg, gx0, gu0, gs = O.adjoint_backward(s["grid"], o, sp, 1.0, s["h"], s["nsteps"], s["x0"], s["u0"], s["kq"],
                                          None, 2 * (uo - s["target"]))
    dev = torch.device("cuda:0")
    g = torch.tensor(s["grid"], device=dev, requires_grad=True)
    u0 = torch.tensor(s["u0"], device=dev, requires_grad=True)
    x0 = torch.tensor(s["x0"], device=dev)
    _, ug = boris_grid_propagate(g, x0, u0, torch.tensor(s["E"], device=dev), O.eplus, [0, 0, 0], [s["sp"]] * 3,
                                 s["h"], s["nsteps"], s["K"])
    ((ug - torch.tensor(s["target"], device=dev)) ** 2).sum().backward()
    assert np.max(np.abs(g.grad.double().cpu().numpy() - gg)) < 1e-6 * np.abs(gg).max()
    assert np.allclose(u0.grad.cpu().numpy(), gu0, rtol=1e-7, atol=1e-8 * np.abs(gu0).max())


def test_adjoint_matches_finite_differences_of_gpu_forward():
    """Central differences of the GPU forward on 256 grid values, 12 initial directions and the scale."""
    s = setup(n=48, nsteps=200, K=16)
    got = gpu_run(s)
    rng = np.random.default_rng(11)
    flat = s["grid"].reshape(-1)
    order = np.argsort(-np.abs(got[3].reshape(-1)))
    picks = np.concatenate([order[:192], rng.choice(order[:2000], 64, replace=False)])
    # central differences converge as e^2: at 1e-3 of the largest value the truncation error is ~3e-3 of
    # the largest gradient, at 1e-5 it is ~3e-8 (measured on the oracle)
    e = np.float32(1e-5) * np.abs(flat).max()
    fd = []
    for j in picks:
        gp = flat.copy()
        gm = flat.copy()
        gp[j] = np.float32(flat[j] + e)
        gm[j] = np.float32(flat[j] - e)
        lp = gpu_run(s, grad=False, grid=gp.reshape(s["grid"].shape))
        lm = gpu_run(s, grad=False, grid=gm.reshape(s["grid"].shape))
        fd.append((lp - lm) / (float(gp[j]) - float(gm[j])))
    fd = np.array(fd)
    ad = got[3].reshape(-1)[picks]
    err = np.abs(ad - fd) / np.abs(got[3]).max()
    big = np.abs(fd) > 1e-2 * np.abs(fd).max()
    rel = np.abs(ad[big] - fd[big]) / np.abs(fd[big])
    print("grid FD: %d values, max err %.2e of the largest gradient; %d large ones: max rel err %.2e"
          % (len(picks), err.max(), big.sum(), rel.max()))
    assert err.max() < 1e-6
    assert rel.max() < 1e-4
    # steps measured on the oracle for this grid: a direction step of 1e-6 leaves a truncation error of
    # 4e-4 relative (the 300-step trajectories are strongly curved), 1e-7 of 1e-7; for the scale,
    # 1e-5 gives 0.16 and 1e-6 gives 7e-9
    for i in range(12):
        for a in range(3):
            d = 1e-7
            up = s["u0"].copy()
            um = s["u0"].copy()
            up[i, a] += d
            um[i, a] -= d
            fdv = (gpu_run(s, grad=False, u0=up) - gpu_run(s, grad=False, u0=um)) / (2 * d)
            assert got[5][i, a] == pytest.approx(fdv, rel=1e-5, abs=1e-6 * np.abs(got[5]).max())
    d = 1e-6
    fds = (gpu_run(s, 1 + d, grad=False) - gpu_run(s, 1 - d, grad=False)) / (2 * d)
    assert got[6] == pytest.approx(fds, rel=1e-6)


def test_adjoint_checkpoint_interval_does_not_change_results():
    s = setup(n=32, nsteps=100, K=7)
    a = gpu_run(s)
    s["K"] = 100
    b = gpu_run(s)
    s["K"] = 1
    c = gpu_run(s)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[2], c[2])
    for r in (a, c):
        assert np.allclose(r[3], b[3], rtol=1e-6, atol=1e-6 * np.abs(b[3]).max())
        assert np.array_equal(r[5], b[5])


def test_adjoint_rejects_bad_input():
    from crpropa3_amd import capi
    from crpropa3_amd.adjoint import boris_grid_propagate
    s = setup(n=4, nsteps=10, K=4)
    dev = torch.device("cuda:0")
    g = torch.tensor(s["grid"], device=dev)
    x0 = torch.tensor(s["x0"], device=dev)
    u0 = torch.tensor(s["u0"], device=dev)
    with pytest.raises(capi.CrpError):
        boris_grid_propagate(g, x0, u0, 1 * O.EeV, O.eplus, [0, 0, 0], [s["sp"]] * 3, s["h"], 10, 0)
    with pytest.raises(ValueError):
        boris_grid_propagate(g.double(), x0, u0, 1 * O.EeV, O.eplus, [0, 0, 0], [s["sp"]] * 3, s["h"], 10, 4)


@pytest.mark.parametrize("budget,shift", [(None, None), ("3e4", None), ("0", None), (None, "4"), ("3e4", "4")])
def test_adjoint_backward_paths(budget, shift, monkeypatch):
    """The backward's three schedules against the oracle on a 20^3 grid (tiles cut by the grid's end):
    the default blocks (17 segments of 8 steps in 6 blocks of at most 3, each recomputed in runs of 8
    steps from sub-checkpoints), one segment per block with the adjoint carried between 17 blocks (a
    small per-block budget), and per-step atomics (budget 0); the blocked ones with 8^3-cell tiles and
    with the 16^3-cell tiles of large grids.  (One block in all: nsteps = K in
    test_adjoint_checkpoint_interval_does_not_change_results.)"""
    if budget is not None:
        monkeypatch.setenv("CRP_ADJOINT_BLOCK_BYTES", budget)
    if shift is not None:
        monkeypatch.setenv("CRP_ADJOINT_TILE_SHIFT", shift)
    s = setup(n=96, N=20, nsteps=130, K=8, seed=9)
    ref = oracle_grad(s)
    got = gpu_run(s)
    rel = np.abs(got[3] - ref[3]).max() / np.abs(ref[3]).max()
    print("budget %s, tile shift %s: grid-gradient max rel %.2e" % (budget, shift, rel))
    assert rel < 1e-6
    assert np.allclose(got[4], ref[4], rtol=1e-7, atol=1e-8 * np.abs(ref[4]).max())
    assert np.allclose(got[5], ref[5], rtol=1e-7, atol=1e-8 * np.abs(ref[5]).max())
    assert got[6] == pytest.approx(ref[6], rel=1e-8)


def test_adjoint_several_particle_groups():
    """600 particles: two full groups of 256 and a partial third one (their tile counts, offsets and
    carried adjoints kept apart), 9 segments of 13 steps (the last of 7) in 5 blocks: runs of 8, 8, 8
    and 2 steps that cross the segments' ends, and one run of 7 in the last block; against the oracle."""
    s = setup(n=600, N=16, nsteps=111, K=13, seed=21)
    ref = oracle_grad(s)
    got = gpu_run(s)
    rel = np.abs(got[3] - ref[3]).max() / np.abs(ref[3]).max()
    print("600 particles: grid-gradient max rel %.2e" % rel)
    assert rel < 1e-6
    assert np.allclose(got[4], ref[4], rtol=1e-7, atol=1e-8 * np.abs(ref[4]).max())
    assert np.allclose(got[5], ref[5], rtol=1e-7, atol=1e-8 * np.abs(ref[5]).max())
    assert got[6] == pytest.approx(ref[6], rel=1e-8)


def test_adjoint_large_grid_uses_16_cell_tiles():
    """A 176^3 grid has 10648 tiles of 8^3 cells, more than the blocked backward takes: it runs with
    16^3-cell tiles (1331) and matches the oracle."""
    s = setup(n=64, N=176, nsteps=60, K=8, seed=12)
    ref = oracle_grad(s)
    got = gpu_run(s)
    rel = np.abs(got[3] - ref[3]).max() / np.abs(ref[3]).max()
    print("176^3: grid-gradient max rel %.2e" % rel)
    assert rel < 1e-6
    assert np.allclose(got[5], ref[5], rtol=1e-7, atol=1e-8 * np.abs(ref[5]).max())
    assert got[6] == pytest.approx(ref[6], rel=1e-8)
