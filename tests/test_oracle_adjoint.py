"""Pins the oracle's FP64 reference gradient of the differentiable Boris model
(orc_adjoint_forward / orc_adjoint_backward).

The forward model is checked against the oracle's restatement of
PropagationBP (src/module/PropagationBP.cpp:19-39, :104-125), which is itself
pinned to the reference's golden trajectories: exactly in a constant grid
(where FP64 and FP32 trilinear interpolation agree) and closely in a noisy
grid.  The hand-written reverse pass is checked against central finite
differences of the forward model for every kind of input (grid values,
positions, directions, field scale, position adjoints) and against the
scaling identity s dL/ds = sum_G G dL/dG."""
import numpy as np
import pytest

import oracle_ctypes as O


def _case(n=6, N=8, nsteps=60, seed=5, amp=1.0):
    sp = 20 * O.kpc
    rng = np.random.default_rng(seed)
    grid = (rng.normal(size=(N, N, N, 3)) * amp * O.nG).astype(np.float32)
    _, E, d = O.sample_particles(seed, n, 1 * O.EeV, 10 * O.EeV)
    u0 = O.normalise(d)
    x0 = rng.uniform(-N * sp, 2 * N * sp, size=(n, 3))  # positions outside the grid wrap periodically
    h = 5 * O.kpc
    kq = O.eplus * O.c_light * h / (2 * E)
    target = O.normalise(rng.normal(size=(n, 3)))
    wx = rng.normal(size=(n, 3)) / (N * sp)
    return dict(grid=grid, origin=np.zeros(3), spacing=np.full(3, sp), h=h, nsteps=nsteps, x0=x0, u0=u0, kq=kq,
                E=E, target=target, wx=wx)


def _loss(c, grid=None, x0=None, u0=None, scale=1.0):
    xo, uo = O.adjoint_forward(c["grid"] if grid is None else grid, c["origin"], c["spacing"], scale, c["h"],
                               c["nsteps"], c["x0"] if x0 is None else x0, c["u0"] if u0 is None else u0, c["kq"])
    return np.sum((uo - c["target"]) ** 2) + np.sum(c["wx"] * xo), xo, uo


def _grads(c, scale=1.0):
    _, xo, uo = _loss(c, scale=scale)
    return O.adjoint_backward(c["grid"], c["origin"], c["spacing"], scale, c["h"], c["nsteps"], c["x0"], c["u0"],
                              c["kq"], c["wx"], 2 * (uo - c["target"]))


def test_forward_equals_boris_in_constant_grid():
    c = _case()
    B = np.array([0.3, -1.1, 0.7]) * O.nG
    c["grid"] = np.broadcast_to(B.astype(np.float32), c["grid"].shape).copy()
    xo, uo = O.adjoint_forward(c["grid"], c["origin"], c["spacing"], 1.0, c["h"], c["nsteps"], c["x0"], c["u0"],
                               c["kq"])
    st = O.OracleState(c["x0"].copy(), c["u0"].copy(), c["E"], np.full(len(c["E"]), O.eplus))
    O.propagate(O.OracleField.uniform(B.astype(np.float32).astype(np.float64)), O.PROP_BP, st, c["nsteps"],
                min_step=c["h"], max_step=c["h"])
    assert np.allclose(xo, st.pos, rtol=1e-13, atol=1e-13 * np.abs(st.pos).max())
    assert np.allclose(uo, st.dir, rtol=0, atol=1e-13)


def test_forward_close_to_boris_in_noisy_grid():
    c = _case(nsteps=20)
    c["x0"] = np.mod(c["x0"], 8 * c["spacing"][0])
    xo, uo = O.adjoint_forward(c["grid"], c["origin"], c["spacing"], 1.0, c["h"], c["nsteps"], c["x0"], c["u0"],
                               c["kq"])
    st = O.OracleState(c["x0"].copy(), c["u0"].copy(), c["E"], np.full(len(c["E"]), O.eplus))
    O.propagate(O.OracleField.grid(c["grid"], c["origin"], c["spacing"]), O.PROP_BP, st, c["nsteps"],
                min_step=c["h"], max_step=c["h"])
    # only the FP32 (reference) versus FP64 (model) interpolation differs
    assert np.max(np.abs(uo - st.dir)) < 1e-6
    assert np.max(np.abs(xo - st.pos)) < 1e-6 * c["h"] * c["nsteps"]


def test_zero_field_is_straight_line_with_identity_jacobian():
    c = _case(nsteps=10)
    c["grid"] = np.zeros_like(c["grid"])
    xo, uo = O.adjoint_forward(c["grid"], c["origin"], c["spacing"], 1.0, c["h"], c["nsteps"], c["x0"], c["u0"],
                               c["kq"])
    assert np.allclose(xo, c["x0"] + c["u0"] * c["h"] * c["nsteps"], rtol=1e-14)
    gx_out = np.random.default_rng(1).normal(size=c["x0"].shape)
    gg, gx0, gu0, gs = O.adjoint_backward(c["grid"], c["origin"], c["spacing"], 1.0, c["h"], c["nsteps"], c["x0"],
                                          c["u0"], c["kq"], gx_out, np.zeros_like(gx_out))
    assert np.allclose(gx0, gx_out, rtol=1e-15)
    # the first step moves along u0 itself, later ones along n(u0):
    # x_final = x0 + h u0 + (L - h) n(u0), so d/du0 = h I + (L - h) (I - u u^T) for |u0| = 1
    L = c["h"] * c["nsteps"]
    u = c["u0"]
    expect = c["h"] * gx_out + (L - c["h"]) * (gx_out - u * np.sum(u * gx_out, axis=1, keepdims=True))
    assert np.allclose(gu0, expect, rtol=1e-9, atol=1e-9 * np.abs(expect).max())
    assert gs == 0.0  # T = 0 everywhere
    # dL/dG at G = 0 is the first-order rotation u x t of the path: for a single step of one particle,
    # x' = x + h u + (h/2) (u x 2t) + O(t^2), so dL/dt = h gx_out x u ... summed over the 8 corners' weights
    assert np.abs(gg).max() > 0


def test_scale_identity():
    c = _case()
    for scale in (1.0, 2.5):
        gg, _, _, gs = _grads(c, scale)
        assert gs * scale == pytest.approx(np.sum(c["grid"].astype(np.float64) * gg), rel=1e-10)


def test_backward_matches_central_differences():
    c = _case()
    gg, gx0, gu0, gs = _grads(c)
    rng = np.random.default_rng(7)
    flat = c["grid"].reshape(-1)
    touched = np.flatnonzero(gg.reshape(-1))
    picks = np.concatenate([touched[np.argsort(-np.abs(gg.reshape(-1)[touched]))[:40]],
                            rng.choice(touched, 20, replace=False)])
    worst = 0.0
    for j in picks:
        e = np.float32(1e-3) * np.abs(flat).max()
        gp = flat.copy()
        gm = flat.copy()
        gp[j] = np.float32(flat[j] + e)
        gm[j] = np.float32(flat[j] - e)
        lp = _loss(c, grid=gp.reshape(c["grid"].shape))[0]
        lm = _loss(c, grid=gm.reshape(c["grid"].shape))[0]
        fd = (lp - lm) / (float(gp[j]) - float(gm[j]))
        worst = max(worst, abs(gg.reshape(-1)[j] - fd) / np.abs(gg).max())
    assert worst < 1e-5
    for i in range(len(c["kq"])):
        for a in range(3):
            for name, arr, g, e in (("u0", c["u0"], gu0, 1e-6), ("x0", c["x0"], gx0, 1e-4 * c["spacing"][0])):
                p = arr.copy()
                m = arr.copy()
                p[i, a] += e
                m[i, a] -= e
                kw_p = {name: p}
                kw_m = {name: m}
                fd = (_loss(c, **kw_p)[0] - _loss(c, **kw_m)[0]) / (2 * e)
                assert g[i, a] == pytest.approx(fd, rel=1e-5, abs=1e-6 * np.abs(g).max()), (name, i, a)
    e = 1e-5
    fd = (_loss(c, scale=1 + e)[0] - _loss(c, scale=1 - e)[0]) / (2 * e)
    assert gs == pytest.approx(fd, rel=1e-6)
