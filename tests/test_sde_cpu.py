"""DiffusionSDE's host interface without a GPU: constructors, defaults, setters and their errors as
the reference's (src/module/DiffusionSDE.cpp:21-46,272-405, test/testDiffusionSDE.py:142-162), the
C-ABI's validation, UniformAdvectionField."""
import ctypes as C

import pytest

import crpropa3_amd as crp
from crpropa3_amd import capi


def test_defaults_and_getters():
    B = crp.UniformMagneticField(crp.Vector3d(0, 0, 1 * crp.nG))
    d = crp.DiffusionSDE(B, 1e-4, 1 * crp.pc, 10 * crp.kpc, 0.)
    assert d.getTolerance() == 1e-4
    assert d.getMinimumStep() == 1 * crp.pc
    assert d.getMaximumStep() == 10 * crp.kpc
    assert d.getEpsilon() == 0.
    assert d.getAlpha() == 1. / 3.
    assert d.getScale() == 1.
    d2 = crp.DiffusionSDE(B)
    assert (d2.getTolerance(), d2.getMinimumStep(), d2.getMaximumStep(), d2.getEpsilon()) == \
        (1e-4, 10 * crp.pc, 1 * crp.kpc, 0.1)
    assert d2.getDescription() == "minStep: 0.01 kpc, maxStep: 1 kpc, tolerance: 0.0001\n"


@pytest.mark.parametrize("call,msg", [
    (lambda d: d.setEpsilon(1.5), "DiffusionSDE: epsilon not in range 0-1"),
    (lambda d: d.setEpsilon(-0.1), "DiffusionSDE: epsilon not in range 0-1"),
    (lambda d: d.setAlpha(2.5), "DiffusionSDE: alpha not in range 0-2"),
    (lambda d: d.setScale(-1), "DiffusionSDE: Scale error: Scale < 0"),
    (lambda d: d.setTolerance(2), "DiffusionSDE: tolerance error not in range 0-1"),
    (lambda d: d.setMinimumStep(-1), "DiffusionSDE: minStep < 0 "),
    (lambda d: d.setMinimumStep(1e30), "DiffusionSDE: minStep > maxStep"),
    (lambda d: d.setMaximumStep(1.0), "DiffusionSDE: maxStep < minStep"),
])
def test_setter_errors(call, msg):
    d = crp.DiffusionSDE(None)
    with pytest.raises(RuntimeError) as e:
        call(d)
    assert str(e.value) == msg


def test_seed_from_random_instance():
    """The device streams' seed is drawn from Random::instance(): seeding it first reproduces it."""
    crp.Random.instance().seed(123)
    s1 = crp.DiffusionSDE(None).getSeed()
    crp.Random.instance().seed(123)
    assert crp.DiffusionSDE(None).getSeed() == s1
    d = crp.DiffusionSDE(None)
    d.setSeed(7)
    assert d.getSeed() == 7


def test_advection_field_and_description():
    a = crp.UniformAdvectionField(crp.Vector3d(1e6, 0, 0))
    assert a.getField(crp.Vector3d(1, 2, 3)).x == 1e6
    assert a.getDivergence(crp.Vector3d(0, 0, 0)) == 0.
    assert a.getDescription() == "v0: 1000 0 0 km/s, "
    d = crp.DiffusionSDE(None, a, 1e-4, 1 * crp.pc, 10 * crp.kpc, 0.2)
    d.setAlpha(0.5)
    d.setScale(2.)
    assert d.getAdvectionFieldAtPosition(crp.Vector3d(0, 0, 0), 0.).x == 1e6
    assert d.getMagneticFieldAtPosition(crp.Vector3d(0, 0, 0), 0.).x == 0
    assert d.getDescription() == ("minStep: 0.001 kpc, maxStep: 10 kpc, tolerance: 0.0001\n"
                                  "epsilon: 0.2, alpha: 0.5\nD_0: 1.22e+25 m^2/s\n")


@pytest.mark.parametrize("args,msg", [
    ((2.0, 1.0, 10.0, 0.1), "DiffusionSDE: tolerance error not in range 0-1"),
    ((1e-4, -1.0, 10.0, 0.1), "DiffusionSDE: minStep < 0 "),
    ((1e-4, 20.0, 10.0, 0.1), "DiffusionSDE: minStep > maxStep"),
    ((1e-4, 1.0, 10.0, 1.1), "DiffusionSDE: epsilon not in range 0-1"),
])
def test_capi_create_sde_validation(args, msg):
    h = C.c_void_p()
    rc = capi.lib().crp_propagator_create_sde(None, *args, C.byref(h))
    assert rc == -1  # CRP_E_INVALID
    assert capi.lib().crp_last_error().decode() == msg


def test_capi_sde_setters():
    h = C.c_void_p()
    assert capi.lib().crp_propagator_create_sde(None, 1e-4, 1.0, 10.0, 0.1, C.byref(h)) == 0
    try:
        assert capi.lib().crp_propagator_set_sde(h, 2.5, 1.0, 0.1, 1) == -1
        assert capi.lib().crp_last_error().decode() == "DiffusionSDE: alpha not in range 0-2"
        assert capi.lib().crp_propagator_set_sde(h, 0.5, -1.0, 0.1, 1) == -1
        assert capi.lib().crp_propagator_set_sde(h, 0.5, 1.0, 0.1, 1) == 0
        v = (C.c_double * 3)(1e6, 0, 0)
        assert capi.lib().crp_propagator_set_sde_advection(h, v) == 0
        assert capi.lib().crp_propagator_set_sde_advection(h, None) == 0
        assert capi.lib().crp_propagator_set_precision(h, capi.PRECISION_MIXED) == -1
    finally:
        capi.lib().crp_propagator_destroy(h)
    # the SDE setters refuse the other propagators
    assert capi.lib().crp_propagator_create_ck(None, 1e-4, 1.0, 10.0, C.byref(h)) == 0
    try:
        assert capi.lib().crp_propagator_set_sde(h, 0.5, 1.0, 0.1, 1) == -1
    finally:
        capi.lib().crp_propagator_destroy(h)
