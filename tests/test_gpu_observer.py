"""Fused Observer(ObserverTimeEvolution) (crp_propagator_set_observer): every
detection's state and the final state against the oracle (correctly rounded
libm) bit for bit, and against the reference's own output where no
transcendental function is on the path (src/module/Observer.cpp:26-57,266-306)."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def grid_field(meta):
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    return capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"])


def run_gpu(c, deactivate, steps, launch=None):
    m = c["meta"]
    a = O.propagator_args(m)
    kind = "BP" if a["prop_type"] == O.PROP_BP else "CK"
    p = capi.Propagator(kind, grid_field(m), a["tolerance"], a["min_step"], a["max_step"],
                        max_length=m.get("maxTrajectoryLength", 0.0))
    p.set_observer(c["times"], deactivate=deactivate)
    if "sphere" in m:
        sp = m["sphere"]
        p.set_spherical_boundary(sp["center"], sp["radius"], sp["margin"], sp["limitStep"])
    if "order" in m:
        p.set_module_order(O.module_args(m)["order"])
    if launch:
        p.set_launch(**launch)
    b = capi.Batch(len(c["energy"]))
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"])
    p.propagate(b, max_steps=steps)
    return b.download(), b.detections(len(c["times"]))


def run_oracle(c, deactivate, steps, libm=O.LIBM_CR):
    m = c["meta"]
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    O.propagate(O.field_for_case(c), steps=steps, max_len=m.get("maxTrajectoryLength", 0.0), libm=libm, state=st,
                observer_times=c["times"], observer_deactivate=deactivate, **O.module_args(m),
                **O.propagator_args(m))
    return st


def same_detections(det, snap):
    ok = ~np.isnan(snap)
    return np.array_equal(ok, ~np.isnan(det)) and bits_equal(det[ok], snap[ok])


@pytest.mark.parametrize("name", ["ck_grid32_observer", "bp_fixed_grid32_observer", "ck_grid32_observer_first",
                                  "bp_fixed_grid32_sphere_observer"])
@pytest.mark.parametrize("deactivate", [False, True])
def test_observer_matches_oracle(name, deactivate):
    c = O.load_case(name)
    steps = int(c["nsteps"].max())
    g, det = run_gpu(c, deactivate, steps)
    st = run_oracle(c, deactivate, steps)
    assert np.array_equal(g["detection_index"], st.det_index)
    assert same_detections(det, st.snap)
    for key in ("pos", "dir", "next_step", "trajectory_length", "time"):
        assert bits_equal(g[key], getattr(st, key)), key
    assert np.array_equal(g["active"], st.active)
    if deactivate:
        assert np.all(g["detection_index"] <= 1) and not g["active"].any()


@pytest.mark.parametrize("name", ["bp_fixed_grid32_observer", "bp_fixed_grid32_sphere_observer"])
def test_observer_fixed_step_bit_exact_vs_reference(name):
    """Fixed-step BP: every detection's state, nextStep as limited by the modules listed before the
    observer (a spherical boundary in the second case), equals the reference's."""
    c = O.load_case(name)
    g, det = run_gpu(c, False, int(c["nsteps"].max()))
    assert np.array_equal(g["detection_index"], c["det_count"])
    for key, col in (("det_pos", slice(0, 3)), ("det_dir", slice(3, 6)), ("det_len", 6), ("det_time", 7),
                     ("det_next", 8)):
        assert same_detections(det[:, :, col], c[key]), key
    assert bits_equal(g["pos"], c["final_pos"]) and bits_equal(g["next_step"], c["final_next"])
    assert np.array_equal(g["active"], c["final_active"])


def test_observer_adaptive_lands_on_the_times():
    """With adaptive steps the observer's limitNextStep makes detections land on their times (to the
    rounding of len + (t - len)), except where a rejected trial shrank the limited step, so that the
    next step overshoots the time: the reference's own run has 23 such detections of 3072 here."""
    c = O.load_case("ck_grid32_observer")
    g, det = run_gpu(c, False, int(c["nsteps"].max()))
    assert np.all(g["detection_index"] == len(c["times"]))
    on_time = np.isclose(det[:, :, 6], c["times"][:, None], rtol=1e-14, atol=0)
    ref_on_time = np.isclose(c["det_len"], c["times"][:, None], rtol=1e-14, atol=0)
    assert np.array_equal(on_time, ref_on_time) and on_time.mean() > 0.99
    # the reference recorded the same lengths
    assert np.allclose(det[:, :, 6], c["det_len"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("launch", [dict(steps_per_launch=13), dict(use_graph=True), dict(block=64)])
def test_observer_independent_of_launch(launch):
    c = O.load_case("ck_grid32_observer")
    steps = int(c["nsteps"].max())
    g0, d0 = run_gpu(c, False, steps)
    g1, d1 = run_gpu(c, False, steps, launch)
    assert np.array_equal(g0["detection_index"], g1["detection_index"])
    assert same_detections(d0, d1)
    assert bits_equal(g0["pos"], g1["pos"])


def test_observer_removed_and_batch_copy():
    c = O.load_case("ck_grid32_observer")
    m = c["meta"]
    a = O.propagator_args(m)
    p = capi.Propagator("CK", grid_field(m), a["tolerance"], a["min_step"], a["max_step"],
                        max_length=m["maxTrajectoryLength"]).set_observer(c["times"])
    b = capi.Batch(len(c["energy"]))
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"])
    p.propagate(b, max_steps=50)
    b2 = capi.Batch(len(c["energy"]))
    b2.copy_from(b)
    assert same_detections(b.detections(12), b2.detections(12))
    assert np.array_equal(b.download()["detection_index"], b2.download()["detection_index"])
    p.set_observer([])
    before = b.download()["detection_index"].copy()
    p.propagate(b, max_steps=50)
    assert np.array_equal(b.download()["detection_index"], before)
    with pytest.raises(capi.CrpError):
        capi.check(capi.lib().crp_propagator_set_observer(p.handle, None, 3, 0))
