"""Host side of the grid turbulence API without a GPU: the reference's spectrum
tests (test/testTurbulentField.cpp:17-50) and its exception cases (:95-113), which
are raised before any device work."""
import re

import pytest

from crpropa3_amd import (GridProperties, Grid3f, Mpc, SimpleGridTurbulence, SimpleTurbulenceSpectrum,
                          TurbulenceSpectrum, Vector3d, initTurbulence, kpc, muG, turbulentCorrelationLength, Gpc)

DEQ = 4 * 2.220446049250313e-16


def test_spectrum_constructor():
    s = TurbulenceSpectrum(1 * muG, 1., 10.)
    assert s.getBrms() == pytest.approx(1 * muG, rel=DEQ)
    assert (s.getLmin(), s.getLmax(), s.getLbendover()) == (1., 10., 1.)
    assert s.getSindex() == pytest.approx(5. / 3., rel=DEQ) and s.getQindex() == 4.


def test_spectrum_correlation_length():
    lBo = 100
    s = TurbulenceSpectrum(1 * muG, 0.00001, 9999999, lBo)
    assert abs(s.getCorrelationLength() - 0.498 * lBo) <= 0.001 * lBo


def test_old_function_for_correlation_length():
    Lc = turbulentCorrelationLength(1 * kpc, 1 * Gpc, -11 / 3.)
    assert abs(Lc - 1 * Gpc / 5) <= 1 * Mpc
    s = SimpleTurbulenceSpectrum(1., 1 * kpc, 1 * Gpc)
    assert s.getCorrelationLength() == pytest.approx(Lc, rel=DEQ)


@pytest.mark.parametrize("lmin,lmax,msg", [(1.5, 8, "turbulentField: lMin < 2 * spacing"),
                                           (8.1, 8, "lMax < lMin"),
                                           (2, 65, "turbulentField: lMax > size")])
def test_turbulence_exceptions(lmin, lmax, msg):
    n = 64
    spacing = 10 * Mpc / n
    grid = Grid3f(Vector3d(0, 0, 0), n, spacing)
    with pytest.raises(RuntimeError, match=re.escape(msg)):
        initTurbulence(grid, 1, lmin * spacing, lmax * spacing)


def test_non_cubic_grid_rejected():
    gp = GridProperties(Vector3d(0.), 16, 16, 8, 1.0)
    with pytest.raises(RuntimeError, match="only cubic grid supported"):
        SimpleGridTurbulence(SimpleTurbulenceSpectrum(1., 2., 8.), gp, 1)
