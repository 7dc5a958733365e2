"""The Python steering layer (pybind11 over the C++ host classes) keeps the
reference's names, constructors, defaults and errors (test/testPropagation.cpp
constructor/exception tests, python/2_headers.i).  No GPU needed."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import (Candidate, EeV, Gpc, Grid3f, GridProperties, MaximumTrajectoryLength, Module,
                          ModuleList, Mpc, ParticleState, PlaneWaveTurbulence, PropagationBP, PropagationCK, Random,
                          SimpleTurbulenceSpectrum, TurbulenceSpectrum, UniformMagneticField, Vector3d, eplus, kpc,
                          muG, nG, nucleusId, pc)


def test_units_match_reference_definitions():
    assert kpc == O.kpc and Mpc == O.Mpc and EeV == O.EeV and nG == O.nG and eplus == O.eplus


@pytest.mark.parametrize("cls", [PropagationBP, PropagationCK])
def test_exceptions(cls):
    """test/testPropagation.cpp:61-81 and :295-315."""
    with pytest.raises(RuntimeError):
        cls(UniformMagneticField(Vector3d(0, 0, 1 * nG)), 0.42, 10, 0)
    with pytest.raises(RuntimeError):
        cls(UniformMagneticField(Vector3d(0, 0, 1 * nG)), 42., 10 * kpc, 20 * kpc)
    propa = cls(UniformMagneticField(Vector3d(0, 0, 1 * nG)))
    propa.setMaximumStep(1 * Mpc)
    with pytest.raises(RuntimeError):
        propa.setTolerance(2.)
    with pytest.raises(RuntimeError):
        propa.setMinimumStep(-1.)
    with pytest.raises(RuntimeError):
        propa.setMinimumStep(2 * Mpc)
    propa.setMinimumStep(0.5 * Mpc)
    with pytest.raises(RuntimeError):
        propa.setMaximumStep(0.1 * Mpc)


@pytest.mark.parametrize("cls", [PropagationBP, PropagationCK])
def test_constructor(cls):
    """test/testPropagation.cpp:84-119 and :318-359."""
    b = UniformMagneticField(Vector3d(0, 0, 1 * nG))
    p = cls(b, 0.01, 1., 100.)
    assert (p.getMinimumStep(), p.getMaximumStep(), p.getTolerance()) == (1., 100., 0.01)
    p.setTolerance(0.0001)
    p.setMinimumStep(10.)
    p.setMaximumStep(10.)
    assert (p.getMinimumStep(), p.getMaximumStep(), p.getTolerance()) == (10., 10., 0.0001)
    if cls is PropagationBP:
        assert PropagationBP(b).getMaximumStep() == 1 * kpc
        assert PropagationBP(b).getTolerance() == 0.42
        fixed = PropagationBP(b, 10 * kpc)
        assert fixed.getMinimumStep() == fixed.getMaximumStep() == 10 * kpc
    else:
        d = PropagationCK(b)
        assert d.getMaximumStep() == 1 * Gpc and d.getMinimumStep() == 0.1 * kpc and d.getTolerance() == 1e-4


def test_descriptions():
    assert "Cash-Karp" in PropagationCK().getDescription()
    assert "Boris" in PropagationBP().getDescription()
    assert "100 Mpc" in MaximumTrajectoryLength(100 * Mpc).getDescription()


def test_particle_state_and_candidate():
    p = ParticleState(nucleusId(56, 26), 10 * EeV, Vector3d(1, 2, 3), Vector3d(0, 3, 4))
    assert p.getCharge() == 26 * eplus
    d = p.getDirection()
    assert (d.x, d.y, d.z) == (0, 0.6, 0.8)
    c = Candidate(p)
    assert c.isActive() and c.getNextStep() == 0 and c.getTrajectoryLength() == 0
    c.setCurrentStep(2.0)
    assert c.getTrajectoryLength() == 2.0 and c.getTime() == 2.0 / O.c_light
    c.setNextStep(5.0)
    c.limitNextStep(3.0)
    assert c.getNextStep() == 3.0
    with pytest.raises(RuntimeError):
        nucleusId(1, 2)


def test_random_matches_reference_sampling():
    """Random restates src/Random.cpp: same draws as the oracle (pinned to the fixtures)."""
    r = Random(44)
    sp, E, d = O.sample_particles(44, 20, 1 * EeV, 300 * EeV, [0.5, 0.2, 0.15, 0.1, 0.05])
    cdf = list(np.cumsum([0.5, 0.2, 0.15, 0.1, 0.05]))
    for i in range(20):
        assert r.randBin(cdf) == sp[i]
        assert r.randPowerLaw(-1, 1 * EeV, 300 * EeV) == E[i]
        v = r.randVector()
        assert (v.x, v.y, v.z) == tuple(d[i])
    assert Random(5489).randInt() == 3499211612


def test_planewave_mode_table_bit_exact():
    c = O.load_case("planewave_c3")
    t = O.planewave_tables(c["meta"])
    pw = PlaneWaveTurbulence(SimpleTurbulenceSpectrum(1 * muG, 1 * pc, 100 * pc), 256, 11)
    assert np.array_equal(np.array(pw.getK()), t["k"])
    assert np.array_equal(np.array(pw.getAk()), t["Ak"])
    assert np.array_equal(np.array([[v.x, v.y, v.z] for v in pw.getKappa()]), t["kappa"])
    g = O.load_case("planewave_general10")
    tg = O.planewave_tables(g["meta"])
    pwg = PlaneWaveTurbulence(TurbulenceSpectrum(1e-4, 1 * pc, 100 * pc), 10, 1)
    assert np.array_equal(np.array(pwg.getAk()), tg["Ak"])
    with pytest.raises(RuntimeError):
        PlaneWaveTurbulence(TurbulenceSpectrum(1e-4, 1 * pc, 100 * pc), 1, 1)


def test_grid_storage_and_numpy_view():
    g = Grid3f(GridProperties(Vector3d(0.), 4, 5, 6, 1.0))
    a = g.array()
    assert a.shape == (4, 5, 6, 3)
    a[1, 2, 3] = [1, 2, 3]
    v = g.getValue(1, 2, 3)
    assert (v.x, v.y, v.z) == (1, 2, 3)
    p = g.positionFromIndex(1 * 30 + 2 * 6 + 3)
    assert (p.x, p.y, p.z) == (1.5, 2.5, 3.5)


def test_modulelist_container():
    ml = ModuleList()
    ml.add(PropagationCK())
    ml.add(MaximumTrajectoryLength(1 * Mpc))
    assert len(ml) == 2
    assert "Cash-Karp" in ml.getDescription()
    ml.remove(1)
    assert len(ml) == 1


def test_python_module_subclass():
    class Counter(Module):
        def __init__(self):
            Module.__init__(self)
            self.n = 0

        def process(self, c):
            self.n += 1

    m = Counter()
    c = Candidate()
    ml = ModuleList()
    ml.add(m)
    ml.process(c)
    assert m.n == 1


def test_observer_time_evolution_lists_and_detection_logic():
    """Host ObserverTimeEvolution: time lists bit-identical to the reference's (fixtures), and
    checkDetection's limitNextStep / DetectionIndex bookkeeping (src/module/Observer.cpp:279-306)."""
    import numpy as np
    import oracle_ctypes as O
    from crpropa3_amd import Candidate, Mpc, Observer, ObserverTimeEvolution, DETECTED, NOTHING
    c = O.load_case("ck_grid32_observer")
    ote = ObserverTimeEvolution(0.5 * Mpc, 10 * Mpc, 12, True)
    assert np.array_equal(np.array(ote.getTimes()).view(np.uint64), c["times"].view(np.uint64))
    b = O.load_case("bp_fixed_grid32_observer")
    assert np.array_equal(np.array(ObserverTimeEvolution(list(b["times"])).getTimes()), b["times"])
    lin = ObserverTimeEvolution(1.0, 2.0, 4)  # min, dist, numb
    assert lin.getTimes() == [1.0, 3.0, 5.0, 7.0]
    ote = ObserverTimeEvolution([1.0, 2.0, 3.0])
    cand = Candidate()
    cand.setNextStep(10.0)
    cand.setCurrentStep(0.4)  # length 0.4 < 1: no detection, nextStep limited to 0.6
    assert ote.checkDetection(cand) == NOTHING and abs(cand.getNextStep() - 0.6) < 1e-15
    cand.setCurrentStep(0.7)  # length 1.1: detection 1, nextStep limited to 2 - 1.1
    cand.setNextStep(10.0)
    assert ote.checkDetection(cand) == DETECTED and cand.getProperty("DetectionIndex") == "1"
    assert abs(cand.getNextStep() - 0.9) < 1e-15
    cand.setCurrentStep(1.0)  # length 2.1: detection 2; index 1 = n - 2: nextStep not limited further
    ns = cand.getNextStep()
    assert ote.checkDetection(cand) == DETECTED and cand.getNextStep() == ns
    obs = Observer()
    obs.add(ObserverTimeEvolution([0.0]))
    cand2 = Candidate()
    obs.process(cand2)
    assert not cand2.isActive()  # Observer deactivates on detection by default


@pytest.mark.parametrize("origin,size,cur,cre,ecur,ecre", [
    ((2, 2, 2), (2, 2, 2), (4.5, 4.3, 4.4), (3, 3, 3), (2.5, 2.3, 2.4), (1, 1, 1)),
    ((0, 0, 0), (2, 2, 2), (-2.5, -0.3, -0.4), (1, 1, 1), (1.5, 1.7, 1.6), (5, 3, 3))])
def test_periodic_box(origin, size, cur, cre, ecur, ecre):
    """test/testBreakCondition.cpp:376-414 (PeriodicBox high / low), host process()."""
    from crpropa3_amd import Candidate, PeriodicBox, Vector3d
    box = PeriodicBox(Vector3d(*origin), Vector3d(*size))
    c = Candidate()
    c.current.setPosition(Vector3d(*cur))
    c.created.setPosition(Vector3d(*cre))
    box.process(c)
    p, q = c.current.getPosition(), c.created.getPosition()
    assert [p.x, p.y, p.z] == pytest.approx(list(ecur), rel=4 * 2.2e-16)
    assert [q.x, q.y, q.z] == pytest.approx(list(ecre), rel=4 * 2.2e-16)


def test_spherical_boundary():
    """test/testBreakCondition.cpp:531-559 (SphericalBoundary inside / outside / limitStep), host process()."""
    from crpropa3_amd import Candidate, SphericalBoundary, Vector3d
    sphere = SphericalBoundary(Vector3d(0, 0, 0), 10)
    c = Candidate()
    c.current.setPosition(Vector3d(9, 0, 0))
    sphere.process(c)
    assert c.isActive() and not c.hasProperty("Rejected")

    sphere = SphericalBoundary(Vector3d(0, 0, 0), 10)
    sphere.setRejectFlag("I passed the galactic border", "Nothing happened")
    c = Candidate()
    c.current.setPosition(Vector3d(0, -10.1, 0))
    sphere.process(c)
    assert not c.isActive() and c.hasProperty("I passed the galactic border")

    sphere = SphericalBoundary(Vector3d(0, 0, 0), 10)
    sphere.setLimitStep(True)
    sphere.setMargin(1)
    c = Candidate()
    c.setNextStep(100)
    c.current.setPosition(Vector3d(0, 0, 9.5))
    sphere.process(c)
    assert c.getNextStep() == pytest.approx(1.5, rel=4 * 2.2e-16)


def test_reject_flag_value_is_the_references():
    """Every condition's default reject flag value is typeid(*this).name() taken in the
    AbstractCondition constructor (src/Module.cpp:20-23): the base's mangled name."""
    from crpropa3_amd import Candidate, MaximumTrajectoryLength, SphericalBoundary, Vector3d
    c = Candidate()
    c.current.setPosition(Vector3d(3, 0, 0))
    SphericalBoundary(Vector3d(0, 0, 0), 1).process(c)
    assert c.getProperty("Rejected") == "N7crpropa17AbstractConditionE"
    c = Candidate()
    c.setTrajectoryLength(2.0)
    MaximumTrajectoryLength(1.0).process(c)
    assert c.getProperty("Rejected") == "N7crpropa17AbstractConditionE"


def test_bench_particles_are_the_reference_draws():
    """bench.py draws its particles with the host Random in the golden fixtures' order (randBin,
    randPowerLaw(-1), randVector per particle), which the oracle's sampler reproduces bit for bit."""
    import bench
    import oracle_ctypes as O
    w = bench.WORKLOADS["ck_grid256"]
    sp, E, d = bench.draw_particles(w["seed"], 4096, w["emin"], w["emax"], w["weights"])
    osp, oE, od = O.sample_particles(w["seed"], 4096, w["emin"], w["emax"], weights=w["weights"])
    assert np.array_equal(sp, osp)
    assert np.array_equal(E.view(np.uint64), oE.view(np.uint64))
    assert np.array_equal(d.view(np.uint64), od.view(np.uint64))
    pos, dn, E2, q = bench.make_particles(w, 4096)
    assert np.array_equal(dn.view(np.uint64), O.normalise(od).view(np.uint64))
    assert set(np.unique(np.round(q / O.eplus))) == {1.0, 2.0, 7.0, 14.0, 26.0}


def test_custom_field_at_position():
    """test/testPythonExtension.py:70-92: getFieldAtPosition of PropagationBP / PropagationCK and
    DiffusionSDE::getMagneticFieldAtPosition forward to the field's getField(position, z)."""
    import crpropa3_amd as crp

    class CustomMagneticField(crp.MagneticField):
        def __init__(self, val):
            crp.MagneticField.__init__(self)
            self.val = val

        def getField(self, position, z):
            return crp.Vector3d(self.val)

    field = CustomMagneticField(crp.gauss)
    propBP = crp.PropagationBP(field, 1e-4, 1 * crp.Mpc, 1 * crp.Mpc)
    propCK = crp.PropagationCK(field, 1e-4, 1 * crp.Mpc, 1 * crp.Mpc)
    propSDE = crp.DiffusionSDE(field)
    pos = crp.Vector3d(-1, 0, 0)
    z = 0
    fieldAtPos = field.getField(pos, z)
    assert fieldAtPos == propBP.getFieldAtPosition(pos, z)
    assert fieldAtPos == propCK.getFieldAtPosition(pos, z)
    assert fieldAtPos == propSDE.getMagneticFieldAtPosition(pos, z)
    # without a field: zero (src/module/PropagationBP.cpp:137-149)
    assert crp.PropagationCK().getFieldAtPosition(pos, z) == crp.Vector3d(0.0)


def test_single_step_api_without_device_field_is_rejected():
    """The single-step functions run on the device: a field defined in Python is rejected with the
    reference's exception type."""
    import crpropa3_amd as crp

    class PyField(crp.MagneticField):
        def __init__(self):
            crp.MagneticField.__init__(self)

        def getField(self, position, z=0):
            return crp.Vector3d(1e-13)

    bp = crp.PropagationBP(PyField(), 1 * crp.kpc)
    with pytest.raises(RuntimeError, match="no GPU implementation"):
        bp.dY(crp.Vector3d(0.0), crp.Vector3d(1, 0, 0), 1 * crp.kpc, 0.0, crp.eplus, crp.EeV)
    assert bp.errorEstimation(crp.Vector3d(3, 0, 0), crp.Vector3d(0, 4, 0), 1.0) == 5 / 0.75


def test_advection_field_classes_match_reference_samples():
    """The reference's advection field classes on the host (src/advectionField/AdvectionField.cpp):
    getField equals the reference's samples bit for bit; a DiffusionSDE accepts every one of them."""
    from crpropa3_amd import (AdvectionFieldList, ConstantSphericalAdvectionField, DiffusionSDE, ObliqueAdvectionShock,
                              OneDimensionalCartesianShock, OneDimensionalSphericalShock, SphericalAdvectionShock,
                              UniformAdvectionField)
    c = O.load_case("advection_fields")
    m = c["meta"]

    (comp, vUp, L), = [par for _, par in m["cartesian_shock"]]
    (rS, vU1, comp1, L1, cool), = [par for _, par in m["spherical_shock_1d_cool"]]
    (compO, vX, vY, LO), = [par for _, par in m["oblique_shock"]]
    (ox, oy, oz, r0, v0, lam, rrot, vphi), = [par for _, par in m["spherical_advection_shock"]]
    sas = SphericalAdvectionShock(Vector3d(ox, oy, oz), r0, v0, lam)
    sas.setRRot(rrot)
    sas.setAzimuthalSpeed(vphi)
    lst = AdvectionFieldList()
    (u,), sph, (cw,) = [par for _, par in m["list"][:1]], m["list"][1][1], [m["list"][2][1]]
    lst.addField(UniformAdvectionField(Vector3d(*u)))
    lst.addField(OneDimensionalSphericalShock(sph[0], sph[1], sph[2], sph[3], bool(sph[4])))
    lst.addField(ConstantSphericalAdvectionField(Vector3d(*cw[:3]), cw[3]))
    fields = {"cartesian_shock": OneDimensionalCartesianShock(comp, vUp, L),
              "spherical_shock_1d_cool": OneDimensionalSphericalShock(rS, vU1, comp1, L1, True),
              "spherical_shock_1d": OneDimensionalSphericalShock(rS, vU1, comp1, L1, False),
              "oblique_shock": ObliqueAdvectionShock(compO, vX, vY, LO),
              "spherical_advection_shock": sas, "list": lst}
    for name, f in fields.items():
        v = np.array([[w.x, w.y, w.z] for w in (f.getField(Vector3d(*p)) for p in c["pos"][:256])])
        assert np.array_equal(v.view(np.uint64), c["v_" + name][:256].view(np.uint64)), name
        DiffusionSDE(UniformMagneticField(Vector3d(0, 0, 1 * nG)), f)
    assert fields["oblique_shock"].getVy() == vY and sas.getRRot() == rrot and fields["spherical_shock_1d_cool"].getCooling()
