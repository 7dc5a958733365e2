"""Seeded turbulent Grid3f values for inputs (benchmarks, tests).

SimpleGridTurbulence::initTurbulence
(src/magneticField/turbulentField/SimpleGridTurbulence.cpp:18-113) run by this
package on the GPU: the k-space field from MT19937 draws in the reference's
order (Random seeded with `seed`), the complex-to-real FFT with hipFFT, the
normalisation to Brms with the reference's rmsFieldStrength/scaleGrid.
"""
import numpy as np


def simple_grid_turbulence(n, spacing, brms, lmin, lmax, seed, sindex=5.0 / 3.0, dtype=np.float32):
    """Return an (n, n, n, 3) array of a Kolmogorov-like turbulent grid with rms |B| = brms."""
    from crpropa3_amd import Grid3f, Vector3d, initTurbulence
    grid = Grid3f(Vector3d(0.0), n, spacing)
    initTurbulence(grid, brms, lmin, lmax, -sindex - 2, seed)
    return np.array(grid.array(), copy=True).astype(dtype, copy=False)
