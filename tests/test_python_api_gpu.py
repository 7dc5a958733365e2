"""The reference's propagation tests (test/testPropagation.cpp,
test/testFunctionalGroups.cpp) through the drop-in Python API, on the GPU."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import (Candidate, EeV, Grid3f, MagneticField, MagneticFieldGrid, MaximumTrajectoryLength,
                          Module, ModuleList, Mpc, ParticleState, PlaneWaveTurbulence, PropagationBP, PropagationCK,
                          TeV, TurbulenceSpectrum, UniformMagneticField, Vector3d, gauss, kpc, nG, nucleusId, pc)

pytestmark = pytest.mark.gpu

DEQ = 4 * 2.220446049250313e-16  # gtest EXPECT_DOUBLE_EQ: within 4 ulp
# the reference's default reject flag value: typeid(*this).name() in the AbstractCondition
# constructor (src/Module.cpp:20-23), the same for every condition
REJECTED = "N7crpropa17AbstractConditionE"


def proton(E=100 * EeV, d=(0, 1, 0)):
    p = ParticleState()
    p.setId(nucleusId(1, 1))
    p.setEnergy(E)
    p.setPosition(Vector3d(0, 0, 0))
    p.setDirection(Vector3d(*d))
    return p


@pytest.mark.parametrize("cls", [PropagationCK, PropagationBP])
def test_zero_field(cls):
    propa = cls(UniformMagneticField(Vector3d(0, 0, 0))) if cls is PropagationCK else \
        PropagationBP(UniformMagneticField(Vector3d(0, 0, 0)), 1 * kpc)
    minStep = 0.1 * kpc
    propa.setMinimumStep(minStep)
    propa.setTolerance(0.42)
    c = Candidate(proton())
    c.setNextStep(0)
    propa.process(c)
    assert c.getCurrentStep() == pytest.approx(minStep, rel=DEQ)
    assert c.getNextStep() == pytest.approx(5 * minStep, rel=DEQ)


@pytest.mark.parametrize("cls", [PropagationCK, PropagationBP])
def test_reduce_step(cls):
    propa = cls(UniformMagneticField(Vector3d(0, 0, 100 * nG)))
    minStep, maxStep = 0.1 * kpc, 1e9 * pc
    propa.setMinimumStep(minStep)
    propa.setMaximumStep(maxStep)
    propa.setTolerance(1e-15)
    c = Candidate(proton(100 * TeV))
    c.setNextStep(maxStep)
    propa.process(c)
    assert c.getCurrentStep() == pytest.approx(minStep, rel=DEQ)
    assert c.getNextStep() == pytest.approx(minStep, rel=DEQ)


@pytest.mark.parametrize("cls", [PropagationCK, PropagationBP])
def test_increase_step(cls):
    propa = cls(UniformMagneticField(Vector3d(0, 0, 1 * nG)))
    minStep, maxStep = 0.001 * pc, 3.125 * pc
    propa.setMinimumStep(minStep)
    propa.setMaximumStep(maxStep)
    propa.setTolerance(0.9)
    c = Candidate(proton())
    for i in range(1, 6):
        propa.process(c)
        assert c.getNextStep() / pc == pytest.approx(minStep * 5 ** i / pc, rel=DEQ)
    assert c.getCurrentStep() / pc == pytest.approx(maxStep / pc / 5., rel=DEQ)
    assert c.getNextStep() / pc == pytest.approx(maxStep / pc, rel=DEQ)


def test_bp_proton():
    propa = PropagationBP(UniformMagneticField(Vector3d(0, 0, 1 * nG)))
    step = 0.01 * kpc
    propa.setMinimumStep(step)
    propa.setMaximumStep(10 * step)
    propa.setTolerance(0.00001)
    c = Candidate(proton())
    c.setNextStep(0)
    propa.process(c)
    assert c.getCurrentStep() == pytest.approx(step, rel=DEQ)
    assert c.getNextStep() == pytest.approx(5 * step, rel=DEQ)


@pytest.mark.parametrize("cls,prec", [(PropagationBP, DEQ), (PropagationCK, 1e-7)])
def test_gyration(cls, prec):
    propa = cls(UniformMagneticField(Vector3d(0, 0, 1 * nG)))
    step = 10. * Mpc
    propa.setMaximumStep(step)
    propa.setMinimumStep(step)
    c = Candidate(proton(d=(1, 1, 1)))
    c.setNextStep(0)
    propa.process(c)
    d, p = c.current.getDirection(), c.current.getPosition()
    assert d.x * d.x + d.y * d.y == pytest.approx(2 / 3., rel=prec)
    assert d.z * d.z == pytest.approx(1 / 3., rel=prec)
    assert p.z * p.z == pytest.approx(step * step / 3., rel=prec)
    for _ in range(9):
        propa.process(c)
    d, p = c.current.getDirection(), c.current.getPosition()
    assert d.x * d.x + d.y * d.y == pytest.approx(2 / 3., rel=prec)
    assert d.z * d.z == pytest.approx(1 / 3., rel=prec)
    assert p.z * p.z == pytest.approx(100 * step * step / 3., rel=prec)


def test_bp_fixed_step_optimization():
    """test/testPropagation.cpp:488-526."""
    fixed = pc
    p1 = PropagationBP(PlaneWaveTurbulence(TurbulenceSpectrum(gauss, pc, 100 * pc), 10, 1), fixed)
    c1 = Candidate(proton(d=(1, 1, 1)))
    for _ in range(9):
        p1.process(c1)
    p2 = PropagationBP(PlaneWaveTurbulence(TurbulenceSpectrum(gauss, pc, 100 * pc), 10, 1), 1, fixed, 1.1 * fixed)
    c2 = Candidate(proton(d=(1, 1, 1)))
    for _ in range(9):
        p2.process(c2)
    for a, b in ((c1.current.getDirection(), c2.current.getDirection()),
                 (c1.current.getPosition(), c2.current.getPosition())):
        assert np.allclose([a.x, a.y, a.z], [b.x, b.y, b.z], rtol=DEQ, atol=0)


@pytest.mark.parametrize("cls,maxs,nxt", [(PropagationCK, 42 * Mpc, 42 * Mpc), (PropagationBP, 1 * kpc, 1 * kpc)])
def test_neutron(cls, maxs, nxt):
    propa = cls(UniformMagneticField(Vector3d(0, 0, 1 * nG)))
    propa.setMinimumStep(1 * kpc)
    propa.setMaximumStep(maxs)
    p = ParticleState()
    p.setId(nucleusId(1, 0))
    p.setEnergy(100 * EeV)
    p.setPosition(Vector3d(0, 0, 0))
    p.setDirection(Vector3d(0, 1, 0))
    c = Candidate(p)
    propa.process(c)
    assert c.getCurrentStep() == pytest.approx(1 * kpc, rel=DEQ)
    assert c.getNextStep() == pytest.approx(nxt, rel=DEQ)
    pos = c.current.getPosition()
    assert (pos.x, pos.y, pos.z) == (0, 1 * kpc, 0)


def test_functional_gyroradius():
    """test/testFunctionalGroups.cpp:10-43: CK through a ModuleList, helix radius E/(qBc)."""
    B = 1 * nG
    E = 100 * EeV
    sim = ModuleList()
    sim.add(PropagationCK(UniformMagneticField(Vector3d(0, 0, B)), 1e-4, 0.1 * kpc, 5 * Mpc))
    sim.add(MaximumTrajectoryLength(2 * np.pi * E / (O.eplus * B * O.c_light)))
    c = Candidate(proton(E, (1, 0, 0)))
    sim.run(c)
    rg = E / (O.eplus * B * O.c_light)
    p = c.current.getPosition()
    assert not c.isActive()
    assert c.getProperty("Rejected") == REJECTED
    assert np.hypot(p.x, p.y + rg) == pytest.approx(rg, rel=1e-3)


def _grid_candidates(n, seed):
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    grid = Grid3f(Vector3d(0.), gm["n"], gm["spacing"])
    grid.array()[:] = g["grid"]
    sp, E, d = O.sample_particles(seed, n, 1 * EeV, 300 * EeV, [0.5, 0.2, 0.15, 0.1, 0.05])
    Z = np.array([1, 2, 7, 14, 26])[sp]
    L = gm["n"] * gm["spacing"]
    cands = [Candidate(nucleusId(60, int(Z[i])), float(E[i]), Vector3d(L / 2, L / 2, L / 2),
                       Vector3d(*d[i])) for i in range(n)]
    return grid, g["grid"], gm, cands, E, Z, d


def test_modulelist_batched_equals_oracle_and_per_candidate():
    grid, gvals, gm, cands, E, Z, d = _grid_candidates(512, 44)
    field = MagneticFieldGrid(grid)
    sim = ModuleList()
    sim.add(PropagationCK(field, 1e-4, 1 * kpc, 1 * Mpc))
    sim.add(MaximumTrajectoryLength(2 * Mpc))
    sim.run(cands)
    assert all(not c.isActive() for c in cands)
    pos = np.array([[c.current.getPosition().x, c.current.getPosition().y, c.current.getPosition().z] for c in cands])
    L = gm["n"] * gm["spacing"]
    st = O.OracleState(np.tile([L / 2] * 3, (512, 1)), d, E, Z * O.eplus)
    O.propagate(O.OracleField.grid(gvals, gm["origin"], gm["spacing"]), O.PROP_CK, st, 10 ** 9, tolerance=1e-4,
                min_step=1 * kpc, max_step=1 * Mpc, max_len=2 * Mpc, libm=O.LIBM_CR)
    assert np.array_equal(pos.view(np.uint64), st.pos.view(np.uint64))
    # the step-wise path (a host module after the propagator) gives the same result
    class Noop(Module):
        def process(self, c):
            pass
    _, _, _, cands2, _, _, _ = _grid_candidates(64, 44)
    sim2 = ModuleList()
    sim2.add(PropagationCK(field, 1e-4, 1 * kpc, 1 * Mpc))
    sim2.add(MaximumTrajectoryLength(2 * Mpc))
    sim2.add(Noop())
    sim2.run(cands2)
    pos2 = np.array([[c.current.getPosition().x, c.current.getPosition().y, c.current.getPosition().z]
                     for c in cands2])
    assert np.array_equal(pos2.view(np.uint64), st.pos[:64].view(np.uint64))


def test_python_field_rejected():
    class MyField(MagneticField):
        def getField(self, pos):
            return Vector3d(0, 0, 1e-13)
    propa = PropagationBP(MyField(), 1 * kpc)
    with pytest.raises(RuntimeError, match="no GPU implementation"):
        propa.process(Candidate(proton()))


def test_grid_interpolate_on_device():
    c = O.load_case("trilinear16")
    m = c["meta"]
    grid = Grid3f(Vector3d(*m["origin"]), 16, m["spacing"])
    grid.array()[:] = c["grid"]
    f = MagneticFieldGrid(grid)
    for i in range(0, 200, 7):
        b = f.getField(Vector3d(*c["pos"][i]))
        assert (b.x, b.y, b.z) == tuple(c["B_periodic"][i])


def test_modulelist_autotuned_and_graph_launches_give_identical_results():
    grid, gvals, gm, cands, E, Z, d = _grid_candidates(512, 45)
    field = MagneticFieldGrid(grid)
    results = []
    for mode in ("default", "autotune", "graph+chunks"):
        _, _, _, cs, _, _, _ = _grid_candidates(512, 45)
        prop = PropagationCK(field, 1e-4, 1 * kpc, 1 * Mpc)
        if mode == "autotune":
            prop.setAutotune(True, 8)
        elif mode == "graph+chunks":
            prop.setLaunchConfig(128, 0, 25, True)
        sim = ModuleList()
        sim.add(prop)
        sim.add(MaximumTrajectoryLength(2 * Mpc))
        sim.run(cs)
        cfg = prop.getLaunchConfig()
        assert cfg[0] in (64, 128, 256) and cfg[4] > 0
        results.append(np.array([[c.current.getPosition().x, c.current.getPosition().y, c.current.getPosition().z]
                                 for c in cs]))
    for r in results[1:]:
        assert np.array_equal(r.view(np.uint64), results[0].view(np.uint64))
    with pytest.raises(RuntimeError):
        PropagationBP(field, 1 * kpc).setLaunchConfig(100)


class _Recorder(Module):
    """Detection action that records (DetectionIndex, position, trajectory length)."""

    def __init__(self):
        Module.__init__(self)
        self.rows = []

    def process(self, c):
        p = c.current.getPosition()
        self.rows.append((int(c.getProperty("DetectionIndex")), p.x, p.y, p.z, c.getTrajectoryLength()))


def _observer_run(extra_module, n=96, seed=51):
    from crpropa3_amd import Observer, ObserverTimeEvolution
    grid, gvals, gm, cands, E, Z, d = _grid_candidates(n, seed)
    ote = ObserverTimeEvolution(0.5 * Mpc, 6 * Mpc, 9, True)
    obs = Observer()
    obs.add(ote)
    obs.setDeactivateOnDetection(False)
    rec = _Recorder()
    obs.onDetection(rec)
    sim = ModuleList()
    sim.add(PropagationCK(MagneticFieldGrid(grid), 1e-4, 1 * kpc, 1 * Mpc))
    sim.add(obs)
    sim.add(MaximumTrajectoryLength(7 * Mpc))
    if extra_module is not None:
        sim.add(extra_module)  # not fusable: every module then runs per step on the host
    sim.run(cands)
    return sorted(rec.rows), cands, np.array(ote.getTimes()), gvals, gm, E, Z, d


def test_modulelist_observer_fused_equals_host_observer_and_oracle():
    """[PropagationCK, Observer(ObserverTimeEvolution), MaximumTrajectoryLength] runs fused on the
    device; with an extra host module it runs step by step with the host Observer.  Both report
    the same detections, equal to the oracle's (kernels' libm)."""
    class Noop(Module):
        def process(self, c):
            pass
    fused, cf, times, gvals, gm, E, Z, d = _observer_run(None)
    stepwise, cs, _, _, _, _, _, _ = _observer_run(Noop())
    assert len(fused) == 9 * len(cf)
    assert fused == stepwise
    L = gm["n"] * gm["spacing"]
    st = O.OracleState(np.tile([L / 2] * 3, (len(E), 1)), d, E, Z * O.eplus)
    O.propagate(O.OracleField.grid(gvals, gm["origin"], gm["spacing"]), O.PROP_CK, st, 10 ** 9, tolerance=1e-4,
                min_step=1 * kpc, max_step=1 * Mpc, max_len=7 * Mpc, libm=O.LIBM_CR, observer_times=times)
    ref = sorted((k + 1, st.snap[k, i, 0], st.snap[k, i, 1], st.snap[k, i, 2], st.snap[k, i, 6])
                 for k in range(len(times)) for i in range(len(E)))
    assert fused == ref
    for cands in (cf, cs):
        pos = np.array([[c.current.getPosition().x, c.current.getPosition().y, c.current.getPosition().z]
                        for c in cands])
        assert np.array_equal(pos.view(np.uint64), st.pos.view(np.uint64))
        assert all(c.getProperty("DetectionIndex") == "9" for c in cands)
        assert all(c.getProperty("Rejected") == REJECTED for c in cands)


def _box_run(extra_module, n=96, seed=53):
    from crpropa3_amd import PeriodicBox
    grid, gvals, gm, cands, E, Z, d = _grid_candidates(n, seed)
    sim = ModuleList()
    sim.add(PropagationCK(MagneticFieldGrid(grid), 1e-4, 1 * kpc, 1 * Mpc))
    sim.add(PeriodicBox(Vector3d(0.1 * Mpc, 0, -0.2 * Mpc), Vector3d(0.7 * Mpc, 0.9 * Mpc, 0.8 * Mpc)))
    sim.add(MaximumTrajectoryLength(5 * Mpc))
    if extra_module is not None:
        sim.add(extra_module)  # not fusable: every module then runs per step on the host
    sim.run(cands)
    pos = np.array([[getattr(getattr(c, s).getPosition(), a) for s in ("current", "previous", "source", "created")
                     for a in "xyz"] for c in cands])
    return pos, gvals, gm, E, Z, d


def test_modulelist_periodic_box_fused_equals_host_and_oracle():
    """[PropagationCK, PeriodicBox, MaximumTrajectoryLength] runs fused on the device; with an extra
    host module the box runs on the host after every step.  Current, previous, source and created
    positions agree bit for bit with each other and with the oracle (kernels' libm)."""
    class Noop(Module):
        def process(self, c):
            pass
    fused, gvals, gm, E, Z, d = _box_run(None)
    stepwise = _box_run(Noop())[0]
    assert np.array_equal(fused.view(np.uint64), stepwise.view(np.uint64))
    L = gm["n"] * gm["spacing"]
    st = O.OracleState(np.tile([L / 2] * 3, (len(E), 1)), d, E, Z * O.eplus)
    box = ((0.1 * Mpc, 0, -0.2 * Mpc), (0.7 * Mpc, 0.9 * Mpc, 0.8 * Mpc))
    O.propagate(O.OracleField.grid(gvals, gm["origin"], gm["spacing"]), O.PROP_CK, st, 10 ** 9, tolerance=1e-4,
                min_step=1 * kpc, max_step=1 * Mpc, max_len=5 * Mpc, libm=O.LIBM_CR, box=box)
    ref = np.hstack([st.pos, st.prev_pos, st.source_pos, st.created_pos])
    assert np.array_equal(fused.view(np.uint64), ref.view(np.uint64))
    lo, hi = np.array(box[0]), np.array(box[0]) + np.array(box[1])
    assert np.all((fused[:, :3] >= lo - 1e-9 * Mpc) & (fused[:, :3] <= hi + 1e-9 * Mpc))


class _FullRecorder(Module):
    """Detection action that records the whole view: index, position, length, nextStep, active,
    reject flag."""

    def __init__(self):
        Module.__init__(self)
        self.rows = []

    def process(self, c):
        p = c.current.getPosition()
        self.rows.append((int(c.getProperty("DetectionIndex")), p.x, p.y, p.z, c.getTrajectoryLength(),
                          c.getNextStep(), c.isActive(), c.getProperty("Rejected") if c.hasProperty("Rejected")
                          else ""))


@pytest.mark.parametrize("order", [("observer", "sphere", "maxlen"), ("sphere", "observer", "maxlen"),
                                   ("maxlen", "sphere", "observer")])
def test_modulelist_sphere_observer_order_fused_equals_host(order):
    """[PropagationCK, then SphericalBoundary, MaximumTrajectoryLength and Observer in the given
    order] runs fused; with an extra host module it runs step by step on the host.  Detection views
    (nextStep, active, reject flag) and final states agree bit for bit, and with the oracle run in
    the same module order."""
    from crpropa3_amd import Observer, ObserverTimeEvolution, SphericalBoundary

    def run(extra):
        grid, gvals, gm, cands, E, Z, d = _grid_candidates(96, 58)
        L = gm["n"] * gm["spacing"]
        times = [0.3 * Mpc, 0.8 * Mpc, 1.5 * Mpc, 2.2 * Mpc, 2.9 * Mpc, 4.0 * Mpc]
        obs = Observer()
        obs.add(ObserverTimeEvolution(times))
        obs.setDeactivateOnDetection(False)
        rec = _FullRecorder()
        obs.onDetection(rec)
        sph = SphericalBoundary(Vector3d(L / 2 + 0.1 * Mpc, L / 2, L / 2), 2.5 * Mpc)
        sph.setMargin(0.03 * Mpc)
        mods = {"observer": obs, "sphere": sph, "maxlen": MaximumTrajectoryLength(3.5 * Mpc)}
        sim = ModuleList()
        sim.add(PropagationCK(MagneticFieldGrid(grid), 1e-4, 1 * kpc, 1 * Mpc))
        for nm in order:
            sim.add(mods[nm])
        if extra is not None:
            sim.add(extra)
        sim.run(cands)
        fin = [(c.current.getPosition().x, c.current.getPosition().y, c.current.getPosition().z, c.getNextStep(),
                c.isActive(), c.getProperty("Rejected")) for c in cands]
        return sorted(rec.rows), fin, gvals, gm, E, Z, d, times, L

    class Noop(Module):
        def process(self, c):
            pass

    fused, ffin, gvals, gm, E, Z, d, times, L = run(None)
    step, sfin = run(Noop())[:2]
    assert fused == step
    assert ffin == sfin
    st = O.OracleState(np.tile([L / 2] * 3, (len(E), 1)), d, E, Z * O.eplus)
    O.propagate(O.OracleField.grid(gvals, gm["origin"], gm["spacing"]), O.PROP_CK, st, 10 ** 9, tolerance=1e-4,
                min_step=1 * kpc, max_step=1 * Mpc, max_len=3.5 * Mpc, libm=O.LIBM_CR, observer_times=times,
                sphere=((L / 2 + 0.1 * Mpc, L / 2, L / 2), 2.5 * Mpc, 0.03 * Mpc, True), order=list(order))
    ref = sorted((k + 1, st.snap[k, i, 0], st.snap[k, i, 1], st.snap[k, i, 2], st.snap[k, i, 6], st.snap[k, i, 8])
                 for k in range(len(times)) for i in range(len(E)) if k < st.det_index[i])
    assert [r[:6] for r in fused] == ref
    assert np.array_equal(np.array([f[:3] for f in ffin]).view(np.uint64), st.pos.view(np.uint64))
    assert all(f[5] == REJECTED for f in ffin)


def test_modulelist_text_output_as_detection_action(tmp_path):
    """A TextOutput (all columns) as the detection action of a fused Observer writes the same lines as
    the step-wise host run (up to line order: fused detections come per candidate)."""
    from crpropa3_amd import Observer, ObserverTimeEvolution, Output, TextOutput

    class Noop(Module):
        def process(self, c):
            pass

    texts = []
    for extra in (None, Noop()):
        grid, gvals, gm, cands, E, Z, d = _grid_candidates(48, 61)
        fn = str(tmp_path / ("out%d.txt" % len(texts)))
        out = TextOutput(fn, Output.Everything)
        obs = Observer()
        obs.add(ObserverTimeEvolution(0.5 * Mpc, 0.5 * Mpc, 6))
        obs.setDeactivateOnDetection(False)
        obs.onDetection(out)
        sim = ModuleList()
        sim.add(PropagationCK(MagneticFieldGrid(grid), 1e-4, 1 * kpc, 1 * Mpc))
        sim.add(obs)
        sim.add(MaximumTrajectoryLength(3.2 * Mpc))
        if extra is not None:
            sim.add(extra)
        sim.run(cands)
        out.close()
        lines = open(fn).read().splitlines()
        body = [ln for ln in lines if not ln.startswith("#")]
        # serial numbers differ between the two runs' candidates: compare the other columns
        texts.append(sorted("\t".join(c for k, c in enumerate(ln.split("\t")) if k not in (3, 12, 21))
                            for ln in body))
        assert len(body) == 6 * 48
    assert texts[0] == texts[1]
