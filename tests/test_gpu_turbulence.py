"""Turbulent grids generated on the GPU (crp_turbulence_generate: k-space field
from MT19937 draws, hipFFT complex-to-real) against the oracle (the same draws,
glibc libm, numpy FFT in double precision), and the reference's own
turbulence tests (test/testTurbulentField.cpp:52-136) through the Python API.

Tolerance: the GPU transforms in FP32 like the reference (fftwf); differences
to the double-precision oracle are FP32 FFT rounding, bounded here by 2e-5 of
the grid's rms value (measured below 2e-6)."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu

TOL = 2e-5


def gpu_grid(n, sp, lmin, lmax, kind, seed, **kw):
    desc = capi.turbulence_desc(n, sp, lmin, lmax, kind, **kw)
    modes = capi.turbulence_mode_count(desc)
    draws = O.randint_batch(seed, modes * (4 if kind == capi.TURBULENCE_SIMPLE else 2))
    return capi.turbulence_generate(desc, draws), modes


@pytest.mark.parametrize("n,lmin_c,lmax_c,kind", [
    (32, 2, 16, capi.TURBULENCE_SIMPLE), (48, 2.5, 48, capi.TURBULENCE_SIMPLE), (33, 2, 11, capi.TURBULENCE_SIMPLE),
    (32, 2, 32, capi.TURBULENCE_GENERAL), (24, 3, 24, capi.TURBULENCE_GENERAL_POWERLAW)])
def test_generator_matches_oracle(n, lmin_c, lmax_c, kind):
    sp, seed = 0.25, 17
    kw = dict(alpha=-11 / 3., lbendover=lmax_c * sp / 6, sindex=5 / 3., qindex=4.0)
    g, modes = gpu_grid(n, sp, lmin_c * sp, lmax_c * sp, kind, seed, **kw)
    ref, rmodes = O.turbulence_grid(n, sp, lmin_c * sp, lmax_c * sp, kind=kind, seed=seed, **kw)
    assert modes == rmodes
    rms = np.sqrt((ref ** 2).sum(-1).mean())
    err = np.abs(g.astype(np.float64) - ref).max() / rms
    assert err < TOL, err


def test_generator_draw_count_is_checked():
    desc = capi.turbulence_desc(16, 1.0, 2.0, 8.0)
    modes = capi.turbulence_mode_count(desc)
    with pytest.raises(capi.CrpError, match="draws"):
        capi.turbulence_generate(desc, np.zeros(4 * modes - 1, dtype=np.uint32))


def test_simple_grid_turbulence_bmean_brms():
    """test/testTurbulentField.cpp:52-71."""
    from crpropa3_amd import GridProperties, Mpc, SimpleGridTurbulence, SimpleTurbulenceSpectrum, Vector3d, \
        meanFieldVector, rmsFieldStrength
    n = 64
    spacing = 10 * Mpc / n
    spectrum = SimpleTurbulenceSpectrum(1, 2 * spacing, 8 * spacing)
    tf = SimpleGridTurbulence(spectrum, GridProperties(Vector3d(0, 0, 0), n, spacing))
    grid = tf.getGrid()
    m = meanFieldVector(grid)
    assert max(abs(m.x), abs(m.y), abs(m.z)) < 1e-7
    assert abs(rmsFieldStrength(grid) - 1) < 1e-7


@pytest.mark.parametrize("general", [False, True])
def test_turbulence_seed(general):
    """test/testTurbulentField.cpp:73-93, 116-136: the same seed gives the same field."""
    from crpropa3_amd import GridProperties, GridTurbulence, Mpc, SimpleGridTurbulence, SimpleTurbulenceSpectrum, \
        TurbulenceSpectrum, Vector3d
    n, spacing = 64, 1 * Mpc
    lmin, lmax = 2 * spacing, 8 * spacing
    if general:
        make = lambda: GridTurbulence(TurbulenceSpectrum(1, lmin, lmax, lmax / 6), GridProperties(Vector3d(0), n,
                                                                                                  spacing), 137)
    else:
        make = lambda: SimpleGridTurbulence(SimpleTurbulenceSpectrum(1, lmin, lmax), GridProperties(Vector3d(0), n,
                                                                                                   spacing), 753)
    tf1, tf2 = make(), make()
    pos = Vector3d(22 * Mpc)
    assert tf1.getField(pos).x == tf2.getField(pos).x
    a = np.array(tf1.getGrid().array(), copy=True)
    b = np.array(tf2.getGrid().array(), copy=True)
    assert np.array_equal(a, b)


def test_init_turbulence_matches_oracle_normalised():
    """initTurbulence(grid, Brms, lMin, lMax, alpha, seed) = the oracle's grid scaled to Brms."""
    from crpropa3_amd import Grid3f, Vector3d, initTurbulence, rmsFieldStrength
    n, sp, seed = 32, 2.0, 21
    grid = Grid3f(Vector3d(0.), n, sp)
    initTurbulence(grid, 3.5, 2 * sp, 16 * sp, -11 / 3., seed)
    g = np.array(grid.array(), copy=True).astype(np.float64)
    ref, _ = O.turbulence_grid(n, sp, 2 * sp, 16 * sp, kind=O.TURB_SIMPLE, alpha=-11 / 3., seed=seed)
    ref *= 3.5 / np.sqrt((ref ** 2).sum(-1).mean())
    assert np.abs(g - ref).max() < TOL * 3.5
    assert abs(rmsFieldStrength(grid) - 3.5) < 1e-6 * 3.5


def test_propagation_through_generated_grid():
    """A SimpleGridTurbulence is a device grid field: PropagationCK runs through it and equals the
    same propagation through a MagneticFieldGrid of its grid."""
    from crpropa3_amd import (Candidate, EeV, GridProperties, MagneticFieldGrid, MaximumTrajectoryLength, ModuleList,
                              Mpc, PropagationCK, SimpleGridTurbulence, SimpleTurbulenceSpectrum, Vector3d, kpc, nG,
                              nucleusId)
    n, sp = 64, 20 * kpc
    tf = SimpleGridTurbulence(SimpleTurbulenceSpectrum(1 * nG, 2 * sp, 32 * sp), GridProperties(Vector3d(0), n, sp),
                              5)
    out = []
    for field in (tf, MagneticFieldGrid(tf.getGrid())):
        cands = [Candidate(nucleusId(1, 1), 10 * EeV, Vector3d(0.5 * Mpc), Vector3d(1, 0.1 * i, 0)) for i in range(8)]
        sim = ModuleList()
        sim.add(PropagationCK(field, 1e-4, 1 * kpc, 1 * Mpc))
        sim.add(MaximumTrajectoryLength(2 * Mpc))
        sim.run(cands)
        out.append([(c.current.getPosition().x, c.current.getPosition().y) for c in cands])
    assert out[0] == out[1]
