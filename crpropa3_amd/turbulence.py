"""Seeded Gaussian turbulent Grid3f generation (input preparation).

Follows the algorithm of SimpleGridTurbulence::initTurbulence
(src/magneticField/turbulentField/SimpleGridTurbulence.cpp:18-113): for every
wave vector of the half spectrum inside [spacing/lMax, spacing/lMin] (in
cycles per cell, K[i] = i/n - floor(i/(n/2)), :41-46) draw a random
orientation perpendicular to k in the basis built from n0 = (1,1,1)
(:70-85), a Gaussian amplitude with sigma = k^(alpha/2) (alpha = -s-2, :89),
a uniform phase (:92-101), transform complex-to-real in 3D
(GridTurbulence.cpp:147-185) and scale the grid to Brms (:112).

FFTW is not available here and the reference only tests this generator
statistically (test/testTurbulentField.cpp), so the values are not bit-for-bit
those of the reference; the same bytes are given to the GPU and to the CPU
checkers.  Random numbers come from numpy's PCG64 seeded with `seed`.
"""
import numpy as np


def check_grid_requirements(n, spacing, lmin, lmax):
    """GridTurbulence::checkGridRequirements (GridTurbulence.cpp:126-143)."""
    if lmin < 2 * spacing:
        raise ValueError("turbulentField: lMin < 2 * spacing")
    if lmax > n * spacing:
        raise ValueError("turbulentField: lMax > size")
    if lmax < lmin:
        raise ValueError("lMax < lMin")


def simple_grid_turbulence(n, spacing, brms, lmin, lmax, seed, sindex=5.0 / 3.0, dtype=np.float32):
    """Return an (n, n, n, 3) float32 grid of Kolmogorov-like turbulence with rms |B| = brms."""
    check_grid_requirements(n, spacing, lmin, lmax)
    alpha = -sindex - 2.0
    n2 = n // 2 + 1
    i = np.arange(n, dtype=np.float64)
    K = i / n - np.floor(i / (n // 2))
    kx = K[:, None, None].astype(np.float32)
    ky = K[None, :, None].astype(np.float32)
    kz = K[None, None, :n2].astype(np.float32)
    shape = (n, n, n2)
    ekx = np.broadcast_to(kx, shape)
    eky = np.broadcast_to(ky, shape)
    ekz = np.broadcast_to(kz, shape)
    k = np.sqrt(ekx * ekx + eky * eky + ekz * ekz)
    kmin = spacing / lmax
    kmax = spacing / lmin
    inside = (k >= kmin) & (k <= kmax)

    # e1 = n0 x ek, e2 = ek x e1 with n0 = (1,1,1); ek parallel to n0 uses a fixed basis
    e1x, e1y, e1z = ekz - eky, ekx - ekz, eky - ekx
    e2x = eky * e1z - e1y * ekz
    e2y = ekz * e1x - e1z * ekx
    e2z = ekx * e1y - e1x * eky
    with np.errstate(invalid="ignore", divide="ignore"):
        cosang = (ekx + eky + ekz) / (k * np.sqrt(np.float32(3.0)))
        parallel = inside & (np.arccos(np.clip(cosang, -1, 1)) < 1e-3)
    e1x = np.where(parallel, -1.0, e1x)
    e1y = np.where(parallel, 1.0, e1y)
    e1z = np.where(parallel, 0.0, e1z)
    e2x = np.where(parallel, 1.0, e2x)
    e2y = np.where(parallel, 1.0, e2y)
    e2z = np.where(parallel, -2.0, e2z)
    with np.errstate(invalid="ignore", divide="ignore"):
        n1 = np.sqrt(e1x * e1x + e1y * e1y + e1z * e1z)
        n2v = np.sqrt(e2x * e2x + e2y * e2y + e2z * e2z)

    rng = np.random.Generator(np.random.PCG64(seed))
    theta = 2 * np.pi * rng.random(shape)
    amp = rng.standard_normal(shape)
    phase = 2 * np.pi * rng.random(shape)
    with np.errstate(invalid="ignore", divide="ignore"):
        scale = np.where(inside, amp * np.power(np.where(inside, k, 1.0).astype(np.float64), alpha / 2), 0.0)
    ph = np.exp(1j * phase)
    ct, st = np.cos(theta), np.sin(theta)
    out = np.empty((n, n, n, 3), dtype=dtype)
    for c, (a1, a2, nn1, nn2) in enumerate(((e1x, e2x, n1, n2v), (e1y, e2y, n1, n2v), (e1z, e2z, n1, n2v))):
        with np.errstate(invalid="ignore", divide="ignore"):
            bc = (a1 / nn1) * ct + (a2 / nn2) * st
        bk = np.where(inside, bc * scale, 0.0) * ph
        out[..., c] = np.fft.irfftn(bk, s=(n, n, n)).astype(dtype)
        del bk
    rms = np.sqrt(np.mean(np.sum(out.astype(np.float64) ** 2, axis=-1)))
    out *= dtype(brms / rms)
    return out


def rms_field_strength(grid):
    """rmsFieldStrength (src/GridTools.cpp:59-69): sqrt(<|B|^2>)."""
    g = np.asarray(grid, dtype=np.float64)
    return float(np.sqrt(np.mean(np.sum(g * g, axis=-1))))


def mean_field_vector(grid):
    """meanFieldVector (src/GridTools.cpp:16-21)."""
    return np.asarray(grid, dtype=np.float64).reshape(-1, 3).mean(axis=0)
