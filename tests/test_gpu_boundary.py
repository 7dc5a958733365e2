"""Fused PeriodicBox and SphericalBoundary (crp_propagator_set_periodic_box,
crp_propagator_set_spherical_boundary): current, previous, source and created
positions, step sizes and rejections after every snapshot, bit for bit against
the reference's output (fixed-step BP, no transcendental on the path) and the
oracle (adaptive CK with the kernels' correctly rounded pow);
src/module/Boundary.cpp:16-27,216-223."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu

KEYS = (("pos", "pos"), ("prev_pos", "previous_pos"), ("source_pos", "source_pos"), ("created_pos", "created_pos"),
        ("dir", "dir"), ("trajectory_length", "trajectory_length"), ("next_step", "next_step"))


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def gpu_run(c, launch=None):
    m = c["meta"]
    a = O.propagator_args(m)
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    field = capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"])
    p = capi.Propagator("BP" if a["prop_type"] == O.PROP_BP else "CK", field, a["tolerance"], a["min_step"],
                        a["max_step"], max_length=m.get("maxTrajectoryLength", 0.0))
    if "box" in m:
        p.set_periodic_box(m["box"]["origin"], m["box"]["size"])
    if "sphere" in m:
        sp = m["sphere"]
        p.set_spherical_boundary(sp["center"], sp["radius"], sp["margin"], sp["limitStep"])
    if launch:
        p.set_launch(**launch)
    b = capi.Batch(len(c["energy"]))
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"])
    done, out = 0, []
    for s in c["snap_steps"]:
        p.propagate(b, max_steps=int(s) - done)
        done = int(s)
        out.append(b.download())
    return out


def test_box_fixed_step_bit_exact_vs_reference():
    c = O.load_case("bp_fixed_grid32_box")
    out = gpu_run(c)
    moved = 0
    for si, g in enumerate(out):
        for key, ref in KEYS:
            assert bits_equal(g[key], c[ref][si]), (si, key)
        assert np.array_equal(g["active"], c["active"][si])
        moved += int(np.sum(np.any(g["source_pos"] != c["init_pos"], axis=1)))
    assert moved > 0  # the box did wrap particles


def test_sphere_fixed_step_bit_exact_vs_reference():
    c = O.load_case("bp_fixed_grid32_sphere")
    out = gpu_run(c)
    for si, g in enumerate(out):
        for key, ref in KEYS:
            assert bits_equal(g[key], c[ref][si]), (si, key)
        assert np.array_equal(g["active"], c["active"][si])
    assert not out[-1]["active"].any()  # the boundary rejected every particle


def test_sphere_removed():
    c = O.load_case("bp_fixed_grid32_sphere")
    m = dict(c["meta"])
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    p = capi.Propagator("BP", capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"]), 1.0, m["fixedStep"],
                        m["fixedStep"])
    sp = m["sphere"]
    p.set_spherical_boundary(sp["center"], sp["radius"], sp["margin"], sp["limitStep"])
    p.set_spherical_boundary(None, 0.0)
    b = capi.Batch(len(c["energy"]))
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"])
    p.propagate(b, max_steps=100)
    assert b.download()["active"].all()


@pytest.mark.parametrize("name", ["ck_grid32_box", "ck_grid32_sphere"])
@pytest.mark.parametrize("launch", [None, dict(steps_per_launch=7, use_graph=True)])
def test_boundary_adaptive_matches_oracle(name, launch):
    c = O.load_case(name)
    m = c["meta"]
    out = gpu_run(c, launch)
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    done = 0
    for si, s in enumerate(c["snap_steps"]):
        O.propagate(O.field_for_case(c), steps=int(s) - done, max_len=m["maxTrajectoryLength"], libm=O.LIBM_CR,
                    state=st, **O.module_args(m), **O.propagator_args(m))
        done = int(s)
        for key, _ in KEYS:
            assert bits_equal(out[si][key], getattr(st, key)), (si, key)
        assert np.array_equal(out[si]["active"], st.active)

