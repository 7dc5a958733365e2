"""Pins the oracle's restatement of the reference's turbulence generators
(src/magneticField/turbulentField/SimpleGridTurbulence.cpp:18-113,
GridTurbulence.cpp:29-124) by inverting its output: every in-range mode's
phase, amplitude and orientation are recovered and compared with what the
MT19937 draws (pinned by the standard known answer) prescribe in the
reference's draw order.  The reference itself needs FFTW, which is absent, so
its grids cannot be produced here: grid values are pinned to the algorithm and
the draws, not to reference output."""
import numpy as np
import pytest

import oracle_ctypes as O

TWO_PI = 2 * np.pi


def k_vectors(n):
    i = np.arange(n)
    K = (i / n - i // (n // 2)).astype(np.float32)
    n2 = n // 2 + 1
    kx, ky, kz = np.meshgrid(K, K, K[:n2], indexing="ij")
    return kx, ky, kz


def inside_modes(n, spacing, lmin, lmax):
    kx, ky, kz = k_vectors(n)
    k = np.sqrt(kx * kx + ky * ky + kz * kz).astype(np.float64)
    return ~((k < spacing / lmax) | (k > spacing / lmin)), kx, ky, kz, k


def unit(v):
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("n,lmin_cells,lmax_cells,seed", [(16, 2, 16, 7), (20, 2.5, 10, 11), (15, 2, 15, 3)])
def test_simple_coefficients_follow_the_draws(n, lmin_cells, lmax_cells, seed):
    sp = 1.0
    alpha = -11 / 3.
    X, modes = O.turbulence_coeffs(n, sp, lmin_cells * sp, lmax_cells * sp, O.TURB_SIMPLE, alpha, seed=seed)
    ins, kx, ky, kz, k = inside_modes(n, sp, lmin_cells * sp, lmax_cells * sp)
    assert modes == ins.sum() and modes > 10
    assert np.all(X[:, ~ins] == 0)
    u = O.randint_batch(seed, 4 * modes).astype(np.float64)
    idx = np.argwhere(ins)  # C order = the reference's loop order (ix, iy, iz)
    for m, (ix, iy, iz) in enumerate(idx):
        B = X[:, ix, iy, iz].astype(np.complex128)
        theta = TWO_PI * u[4 * m] / 4294967295.0
        g = np.sqrt(-2.0 * np.log(1.0 - (u[4 * m + 1] + 0.5) / 4294967296.0)) * \
            np.cos(TWO_PI * u[4 * m + 2] / 4294967296.0)
        phase = TWO_PI * u[4 * m + 3] / 4294967295.0
        b = (B * np.exp(-1j * phase))
        assert np.all(np.abs(b.imag) <= 1e-6 * np.abs(b).max() + 1e-30)  # one common phase
        b = b.real
        kk = k[ix, iy, iz]
        assert np.linalg.norm(b) == pytest.approx(abs(g) * kk ** (alpha / 2), rel=1e-5)
        ek = np.array([kx[ix, iy, iz], ky[ix, iy, iz], kz[ix, iy, iz]], dtype=np.float64)
        assert abs(np.dot(b, ek)) <= 1e-5 * np.linalg.norm(b) * np.linalg.norm(ek)  # transverse
        # orientation: b / g = k^(alpha/2) (e1 cos theta + e2 sin theta)
        if np.degrees(np.arccos(np.clip(ek.sum() / np.sqrt(3) / np.linalg.norm(ek), -1, 1))) < 1e-3 * 180 / np.pi:
            e1, e2 = unit(np.array([-1., 1, 0])), unit(np.array([1., 1, -2]))
        else:
            e1 = unit(np.cross([1., 1, 1], ek))
            e2 = unit(np.cross(ek, e1))
        d = b / (g * kk ** (alpha / 2))
        assert d @ e1 == pytest.approx(np.cos(theta), abs=2e-5)
        assert d @ e2 == pytest.approx(np.sin(theta), abs=2e-5)


@pytest.mark.parametrize("kind", [O.TURB_GENERAL, O.TURB_GENERAL_POWERLAW])
def test_general_coefficients_follow_the_spectrum(kind):
    n, sp, seed = 16, 0.5, 5
    lmin, lmax, lbo, s, q = 2 * sp, 16 * sp, 8 * sp / 6, 5 / 3., 4.0
    X, modes = O.turbulence_coeffs(n, sp, lmin, lmax, kind, lbendover=lbo, sindex=s, qindex=q, seed=seed)
    ins, kx, ky, kz, k = inside_modes(n, sp, lmin, lmax)
    assert modes == ins.sum()
    u = O.randint_batch(seed, 2 * modes).astype(np.float64)
    lam = 1 / sp * 2 * np.pi
    for m, (ix, iy, iz) in enumerate(np.argwhere(ins)):
        B = X[:, ix, iy, iz].astype(np.complex128)
        phase = TWO_PI * u[2 * m + 1] / 4294967295.0
        b = (B * np.exp(-1j * phase)).real
        kk = k[ix, iy, iz] * lam
        if kind == O.TURB_GENERAL:
            E = (kk * lbo) ** q / (1 + (kk * lbo) ** 2) ** ((s + q) / 2 + 1)
        else:
            E = kk ** (-s - 2)
        assert np.linalg.norm(b) == pytest.approx(np.sqrt(E), rel=1e-5)  # deterministic amplitude


def test_oracle_grid_mean_zero_and_divergence_free():
    """Real-space field of the oracle: zero mean (k = 0 is outside the range) and zero spectral
    divergence (every mode is transverse)."""
    n, sp = 32, 1.0
    g, modes = O.turbulence_grid(n, sp, 2 * sp, 16 * sp, kind=O.TURB_SIMPLE, seed=9)
    rms = np.sqrt((g ** 2).sum(-1).mean())
    assert np.abs(g.mean(axis=(0, 1, 2))).max() <= 1e-12 * rms
    F = [np.fft.fftn(g[..., c]) for c in range(3)]
    f = np.fft.fftfreq(n)
    kx, ky, kz = np.meshgrid(f, f, f, indexing="ij")
    div = kx * F[0] + ky * F[1] + kz * F[2]
    assert np.abs(div).max() <= 1e-5 * np.abs(F[0]).max()
