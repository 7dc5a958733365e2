"""DiffusionSDE through the reference's Python interface on the GPU: the reference's unit tests
(test/testDiffusionSDE.py:165-233: neutral, no field, advection) and ModuleList runs, whose random
numbers depend only on the seed, the candidates' serial numbers and their step counts -- not on
how the run is split -- and equal the oracle's on the same streams."""
import numpy as np
import pytest

import crpropa3_amd as crp
import oracle_ctypes as O

pytestmark = pytest.mark.gpu


def candidate(Z, E, direction=(1, 0, 0)):
    c = crp.Candidate()
    c.current.setId(crp.nucleusId(1, Z))
    c.current.setEnergy(E)
    c.current.setDirection(crp.Vector3d(*direction))
    return c


def test_reference_neutral_propagation():
    B = crp.UniformMagneticField(crp.Vector3d(0, 0, 1 * crp.nG))
    d = crp.DiffusionSDE(B, 1e-4, 1 * crp.pc, 10 * crp.kpc, 0.)
    c = candidate(0, 10 * crp.TeV)
    d.process(c)
    pos = c.current.getPosition()
    assert pos.x / crp.pc == pytest.approx(1., abs=1e-7)
    assert pos.y == 0. and pos.z == 0.
    assert c.getNextStep() / (10 * crp.kpc) == pytest.approx(1., abs=1e-7)
    assert c.getDiffusionSteps() == 1


def test_reference_no_field_propagation():
    d = crp.DiffusionSDE(crp.UniformMagneticField(crp.Vector3d(0.)), 1e-4, 1 * crp.pc, 10 * crp.kpc, 0.)
    c = candidate(1, 10 * crp.TeV)
    d.process(c)
    pos = c.current.getPosition()
    assert pos.x / crp.pc == pytest.approx(1., abs=1e-7)
    assert pos.y == 0. and pos.z == 0.
    assert c.getNextStep() / (1 * crp.pc) == pytest.approx(5., abs=1e-7)


def test_reference_advective_propagation():
    adv = crp.UniformAdvectionField(crp.Vector3d(0., 1e6, 0.))
    d = crp.DiffusionSDE(crp.UniformMagneticField(crp.Vector3d(0.)), adv, 1e-4, 1 * crp.pc, 10 * crp.kpc, 0.)
    c = candidate(1, 10 * crp.TeV)
    d.process(c)
    pos = c.current.getPosition()
    assert pos.x == 0.
    assert pos.y == pytest.approx(1 * crp.pc / crp.c_light * 1e6, abs=1e-7)
    assert pos.z == 0.
    assert c.getNextStep() / (1 * crp.pc) == pytest.approx(5., abs=1e-7)


def make_candidates(n, seed=3, serials=None):
    """n candidates (Z 0..2, 1-100 PeV) from a seeded generator; serials: their serial numbers"""
    rng = np.random.default_rng(seed)
    cands = []
    for i in range(n):
        Z = int(rng.integers(0, 3))
        c = crp.Candidate(crp.nucleusId(max(1, 2 * Z), Z), float(rng.uniform(1, 100)) * crp.PeV,
                          crp.Vector3d(*rng.uniform(0, 1, 3) * crp.kpc), crp.Vector3d(*rng.normal(size=3)))
        if serials is not None:
            c.setSerialNumber(int(serials[i]))
        cands.append(c)
    return cands


def state(cands):
    pos = np.array([[c.current.getPosition().x, c.current.getPosition().y, c.current.getPosition().z] for c in cands])
    d = np.array([[c.current.getDirection().x, c.current.getDirection().y, c.current.getDirection().z] for c in cands])
    length = np.array([c.getTrajectoryLength() for c in cands])
    nxt = np.array([c.getNextStep() for c in cands])
    steps = np.array([c.getDiffusionSteps() for c in cands])
    return pos, d, length, nxt, steps


ADVECTIONS = {
    "uniform": ([0, 0, 5e5], lambda: crp.UniformAdvectionField(crp.Vector3d(0, 0, 5e5))),
    "cspherical": ({"kind": 2, "origin": [0.5 * O.kpc, 0, 0], "par": [3e5]},
                   lambda: crp.ConstantSphericalAdvectionField(crp.Vector3d(0.5 * O.kpc, 0, 0), 3e5)),
    "spherical": ({"kind": 3, "origin": [0.5 * O.kpc, 0, 0], "par": [8 * O.kpc, 4e5, 2 * O.kpc, 1.5]},
                  lambda: crp.SphericalAdvectionField(crp.Vector3d(0.5 * O.kpc, 0, 0), 8 * O.kpc, 4e5, 2 * O.kpc, 1.5)),
    "shocks": ({"terms": [(O.ADV_CARTESIAN_SHOCK, [4., 1e6, 0.3 * O.kpc]),
                          (O.ADV_SPHERICAL_SHOCK_1D, [1 * O.kpc, 8e5, 3., 0.4 * O.kpc, 1]),
                          (O.ADV_OBLIQUE_SHOCK, [3., 5e5, 2e5, 0.2 * O.kpc]),
                          (O.ADV_SPHERICAL_SHOCK, [0.5 * O.kpc, 0, 0, 1 * O.kpc, 6e5, 0.3 * O.kpc, 2 * O.kpc, 2e5])]},
               lambda: _shock_list()),
}


def _shock_list():
    """AdvectionFieldList of the four shock fields (src/advectionField/AdvectionField.cpp:182-532)"""
    lst = crp.AdvectionFieldList()
    lst.addField(crp.OneDimensionalCartesianShock(4., 1e6, 0.3 * crp.kpc))
    lst.addField(crp.OneDimensionalSphericalShock(1 * crp.kpc, 8e5, 3., 0.4 * crp.kpc, True))
    lst.addField(crp.ObliqueAdvectionShock(3., 5e5, 2e5, 0.2 * crp.kpc))
    sas = crp.SphericalAdvectionShock(crp.Vector3d(0.5 * crp.kpc, 0, 0), 1 * crp.kpc, 6e5, 0.3 * crp.kpc)
    sas.setRRot(2 * crp.kpc)
    sas.setAzimuthalSpeed(2e5)
    lst.addField(sas)
    return lst


def sim_for(seed, adv="uniform"):
    B = crp.UniformMagneticField(crp.Vector3d(1e-13, 0, 3e-13))
    d = crp.DiffusionSDE(B, ADVECTIONS[adv][1](), 1e-4, 1 * crp.pc, 200 * crp.pc, 0.3)
    d.setSeed(seed)
    sim = crp.ModuleList()
    sim.add(d)
    sim.add(crp.MaximumTrajectoryLength(20 * crp.kpc))
    return sim, d


def test_modulelist_split_runs_and_process_loop_agree():
    """One run to the end, the same candidates in runs of a few steps, and process() calls of the
    module one candidate at a time give the same numbers (same serial numbers, same seed)."""
    n = 24
    a = make_candidates(n)
    serials = [c.getSerialNumber() for c in a]
    b = make_candidates(n, serials=serials)
    cc = make_candidates(n, serials=serials)
    sim, d = sim_for(77)
    sim.run(a)
    sim2, _ = sim_for(77)
    sim2.setMaximumSteps(5)
    for _ in range(200):
        sim2.run(b)
    sim3, d3 = sim_for(77)
    maxlen = crp.MaximumTrajectoryLength(20 * crp.kpc)
    for c in cc:
        while c.isActive():
            d3.process(c)
            maxlen.process(c)
    sa, sb, sc = state(a), state(b), state(cc)
    for x, y, z in zip(sa, sb, sc):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    assert not any(c.isActive() for c in a + b)


@pytest.mark.parametrize("adv", sorted(ADVECTIONS))
def test_modulelist_matches_oracle_streams(adv):
    n = 48
    cands = make_candidates(n, seed=9)
    pos0, dir0, _, _, _ = state(cands)
    E = np.array([c.current.getEnergy() for c in cands])
    q = np.array([c.current.getCharge() for c in cands])
    serial = np.array([c.getSerialNumber() for c in cands], dtype=np.uint64)
    sim, d = sim_for(5, adv)
    sim.run(cands)
    st = O.OracleState(pos0, dir0, E, q)
    st.dir = dir0.copy()  # the candidates' unit vectors as they are (normalising again may change bits)
    steps = np.zeros(n, dtype=np.uint32)
    O.sde_propagate(O.OracleField.uniform([1e-13, 0, 3e-13]), st, 10 ** 6, tolerance=1e-4, min_step=1 * crp.pc,
                    max_step=200 * crp.pc, epsilon=0.3, advection=ADVECTIONS[adv][0], max_len=20 * crp.kpc,
                    rng=O.SDE_PHILOX, seed=5, sde_step=steps, sde_stream=serial)
    pos, dd, length, nxt, k = state(cands)
    assert np.array_equal(pos.view(np.uint64), st.pos.view(np.uint64))
    assert np.array_equal(dd.view(np.uint64), st.dir.view(np.uint64))
    assert np.array_equal(length.view(np.uint64), st.trajectory_length.view(np.uint64))
    assert np.array_equal(k, steps)


class _Noop(crp.Module):
    """a host module after the others: the list no longer fuses and runs step by step"""

    def process(self, c):
        pass


class _Recorder(crp.Module):
    """detection action: the candidate's state at detection (the step-by-step path hands over the live
    candidate, as the reference does, so a collector would show later states)"""

    def __init__(self):
        crp.Module.__init__(self)
        self.rows = []

    def process(self, c):
        p = c.current.getPosition()
        self.rows.append((c.getSerialNumber(), c.getTrajectoryLength(), p.x, p.y, p.z, c.getNextStep()))


def _fused_or_stepwise(stepwise, serials=None):
    cands = make_candidates(40, seed=12, serials=serials)
    B = crp.UniformMagneticField(crp.Vector3d(1e-13, 2e-13, 3e-13))
    d = crp.DiffusionSDE(B, 1e-4, 1 * crp.pc, 100 * crp.pc, 0.2)
    d.setSeed(3)
    rec = _Recorder()
    obs = crp.Observer()
    obs.add(crp.ObserverTimeEvolution(0.5 * crp.kpc, 4 * crp.kpc, 6))
    obs.setDeactivateOnDetection(False)
    obs.onDetection(rec)
    sim = crp.ModuleList()
    sim.add(d)
    sim.add(crp.PeriodicBox(crp.Vector3d(-2 * crp.kpc), crp.Vector3d(4 * crp.kpc)))
    sim.add(crp.SphericalBoundary(crp.Vector3d(0.), 1.5 * crp.kpc))
    sim.add(obs)
    sim.add(crp.MaximumTrajectoryLength(5 * crp.kpc))
    if stepwise:
        sim.add(_Noop())
    sim.run(cands)
    return cands, sorted(rec.rows)


def test_modulelist_sde_with_box_sphere_observer_fused_equals_stepwise():
    """[DiffusionSDE, PeriodicBox, SphericalBoundary, Observer(ObserverTimeEvolution), MaximumTrajectoryLength]
    in one device run equals the same list run step by step with the host modules: positions, lengths,
    flags, step counts and every detection"""
    a, det_a = _fused_or_stepwise(False)
    b, det_b = _fused_or_stepwise(True, serials=[c.getSerialNumber() for c in a])
    sa, sb = state(a), state(b)
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    assert [c.isActive() for c in a] == [c.isActive() for c in b]
    flags = [(c.hasProperty("Rejected"), c.hasProperty("DetectionIndex") and c.getProperty("DetectionIndex"))
             for c in a]
    assert flags == [(c.hasProperty("Rejected"), c.hasProperty("DetectionIndex") and c.getProperty("DetectionIndex"))
                     for c in b]
    assert len(det_a) > 20 and det_a == det_b
    assert any(not c.isActive() and c.getTrajectoryLength() < 5 * crp.kpc for c in a)  # the sphere rejected some


def test_substep_limit_keeps_every_candidates_progress(monkeypatch):
    """A DiffusionSDE step whose trials would need more sub-steps than the limit stops its candidate
    where it is (inactive) and the call raises, after every candidate's device state is back on the
    host.  The limit (2^20 sub-steps in production) is lowered to 2^6 from the host
    (CRP_SDE_MAX_HALVINGS) so that a 10 pc first step in a 1 pc white-noise grid at a 1e-8 m tolerance
    (about 2^15-2^17 sub-steps without the limit) reaches it, while a candidate starting with 1e-12 pc
    steps (one sub-step each) runs to its length limit untouched."""
    monkeypatch.setenv("CRP_SDE_MAX_HALVINGS", "6")
    rng = np.random.default_rng(4)
    g = crp.Grid3f(crp.Vector3d(0.), 16, 1 * crp.pc)
    g.array()[:] = (rng.normal(size=(16, 16, 16, 3)) * crp.nG).astype(np.float32)
    d = crp.DiffusionSDE(crp.MagneticFieldGrid(g), 1e-8, 1e-20 * crp.pc, 10 * crp.pc, 0.1)
    d.setSeed(3)
    sim = crp.ModuleList()
    sim.add(d)
    sim.add(crp.MaximumTrajectoryLength(1e-9 * crp.pc))
    start = crp.Vector3d(8 * crp.pc, 8 * crp.pc, 8 * crp.pc)
    cands = []
    for serial, nxt in ((101, 1e-12 * crp.pc), (102, 10 * crp.pc)):
        c = crp.Candidate(crp.nucleusId(1, 1), 1 * crp.PeV, start, crp.Vector3d(1, 0, 0))
        c.setSerialNumber(serial)  # names the candidate's random stream
        c.setNextStep(nxt)
        cands.append(c)
    with pytest.raises(RuntimeError, match="sub-steps"):
        sim.run(cands)
    small, big = cands
    print("small: length %.3e pc, %d steps; big: length %.3e pc, %d steps" % (
        small.getTrajectoryLength() / crp.pc, small.getDiffusionSteps(), big.getTrajectoryLength() / crp.pc,
        big.getDiffusionSteps()))
    assert not big.isActive()
    # stopped where it was; its one step's random numbers are counted as drawn
    assert big.getTrajectoryLength() == 0.0 and big.getDiffusionSteps() == 1
    p = big.current.getPosition()
    assert (p.x, p.y, p.z) == (start.x, start.y, start.z)
    assert not small.isActive()
    assert small.getTrajectoryLength() == pytest.approx(1e-9 * crp.pc, rel=1e-9)
    assert small.getDiffusionSteps() > 1
    # one process() call of the capped case: the same error, the candidate stopped in place
    c = crp.Candidate(crp.nucleusId(1, 1), 1 * crp.PeV, start, crp.Vector3d(1, 0, 0))
    c.setSerialNumber(102)
    c.setNextStep(10 * crp.pc)
    with pytest.raises(RuntimeError, match="sub-steps"):
        d.process(c)
    assert not c.isActive() and c.getTrajectoryLength() == 0.0
