"""Tricubic and nearest-neighbour grids on the GPU (crp_field_set_grid_interpolation; Grid::interpolate,
include/crpropa/Grid.h:318-335, 383-536): field samples and fixed-step trajectories bit for bit against
the output of the reference's SIMD_EXTENSIONS=avx build, adaptive trajectories and decorated fields bit
for bit against the oracle (which matches that output too), and the device-layout rules of such grids."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu

NAMES = {O.GRID_TRILINEAR: "TRILINEAR", O.GRID_TRICUBIC: "TRICUBIC", O.GRID_NEAREST: "NEAREST_NEIGHBOUR"}
MODES = [("tricubic_periodic", O.GRID_TRICUBIC, False, False),
         ("tricubic_reflective", O.GRID_TRICUBIC, True, False),
         ("tricubic_clip", O.GRID_TRICUBIC, False, True),
         ("nearest_periodic", O.GRID_NEAREST, False, False),
         ("nearest_reflective", O.GRID_NEAREST, True, False)]


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("mode,ip,refl,clip", MODES)
def test_grid3f_samples_bit_exact_vs_reference(mode, ip, refl, clip):
    c = O.load_case("interp_grids")
    m = c["meta"]
    f = capi.Field.grid3f(c["grid"], m["origin"], m["spacing"], reflective=refl, clip_volume=clip)
    f.set_interpolation(NAMES[ip])
    assert f.interpolation() == NAMES[ip] and f.layout() == "points"
    B = f.eval(c["pos"])
    assert bits_equal(B, c["B_" + mode])
    # the mixed-precision path evaluates these grids with the FP64 path's code (an FP32 result)
    assert bits_equal(f.eval(c["pos"], capi.PRECISION_MIXED), B)


@pytest.mark.parametrize("mode,ip,refl,clip", MODES)
def test_grid1f_samples_bit_exact_vs_reference(mode, ip, refl, clip):
    """The Grid1f of a modulated field, read through a one-point Grid3f of value (1, 1, 1) taken
    nearest-neighbour, so that B = m exactly."""
    c = O.load_case("interp_grids")
    m = c["meta"]
    ones = np.ones((1, 1, 1, 3), np.float32)
    f = capi.Field.modulated_grid3f(ones, [0, 0, 0], [1, 1, 1], c["mod"], m["mod_origin"], m["mod_spacing"],
                                    mreflective=refl, mclip_volume=clip)
    f.set_interpolation("NEAREST_NEIGHBOUR", "values").set_interpolation(NAMES[ip], "modulation")
    B = f.eval(c["pos"])
    want = c["m_" + mode].astype(np.float64)
    for a in range(3):
        assert bits_equal(B[:, a], want), a
    assert bits_equal(f.eval(c["pos"], capi.PRECISION_MIXED), B)


def run_gpu(c, precision=capi.PRECISION_FP64, launch=None):
    m = c["meta"]
    a = O.propagator_args(m)
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    if m["field"] == "grid32_modulated":
        field = capi.Field.modulated_grid3f(g["grid"], gm["origin"], gm["spacing"], c["mod"], m["mod_origin"],
                                            m["mod_spacing"], mreflective=m["mod_reflective"])
        field.set_interpolation(m["interpolation"], "modulation")
    else:
        field = capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"], reflective=m.get("reflective", False))
    field.set_interpolation(m["interpolation"])
    p = capi.Propagator("BP" if a["prop_type"] == O.PROP_BP else "CK", field, a["tolerance"], a["min_step"],
                        a["max_step"], max_length=m.get("maxTrajectoryLength", 0.0), precision=precision)
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


@pytest.mark.parametrize("name", ["bp_fixed_grid32_tricubic_reflective", "bp_fixed_grid32_nearest_reflective"])
def test_fixed_step_bit_exact_vs_reference(name):
    c = O.load_case(name)
    for si, g in enumerate(run_gpu(c)):
        assert bits_equal(g["pos"], c["pos"][si]) and bits_equal(g["dir"], c["dir"][si]), si


@pytest.mark.parametrize("launch", [None, dict(steps_per_launch=7, use_graph=True)])
@pytest.mark.parametrize("name", ["ck_grid32_tricubic", "ck_grid32_modulated_tricubic"])
def test_adaptive_matches_oracle(name, launch):
    c = O.load_case(name)
    m = c["meta"]
    out = run_gpu(c, launch=launch)
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    done = 0
    for si, s in enumerate(c["snap_steps"]):
        O.propagate(O.field_for_case(c), steps=int(s) - done, max_len=m["maxTrajectoryLength"], libm=O.LIBM_CR,
                    state=st, **O.propagator_args(m))
        done = int(s)
        for key in ("pos", "dir", "next_step", "trajectory_length"):
            assert bits_equal(out[si][key], getattr(st, key)), (si, key)
        assert np.array_equal(out[si]["active"], st.active)


def test_mixed_precision_tricubic_fixed_step_close():
    """The mixed path (FP32 rotation, FP64 drift) over a tricubic grid stays near the FP64 path."""
    c = O.load_case("bp_fixed_grid32_tricubic_reflective")
    ref = run_gpu(c)
    mix = run_gpu(c, precision=capi.PRECISION_MIXED)
    step = c["meta"]["fixedStep"]
    assert np.abs(mix[1]["pos"] - ref[1]["pos"]).max() <= 1e-5 * step * 10
    assert np.abs(mix[1]["pos"] - ref[1]["pos"]).max() > 0


def test_layout_and_storage_follow_interpolation():
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    f = capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"])
    assert f.layout() == "cells" and f.interpolation() == "TRILINEAR"
    f.set_storage("fp16")
    f.set_interpolation("TRICUBIC")
    assert f.layout() == "points" and f.interpolation() == "TRICUBIC"
    with pytest.raises(capi.CrpError, match="trilinear"):
        f.set_layout("cells")
    with pytest.raises(capi.CrpError, match="trilinear"):
        f.set_storage("fp16")
    with pytest.raises(capi.CrpError, match="no such grid"):
        f.set_interpolation("TRICUBIC", "modulation")
    assert capi.lib().crp_field_set_grid_interpolation(f.handle, 0, 7) != 0
    assert "unknown interpolation type" in capi.lib().crp_last_error().decode()
    f.set_interpolation("NEAREST_NEIGHBOUR")
    f.set_interpolation("TRILINEAR")
    assert f.layout() == "cells"
    pos = np.random.default_rng(1).uniform(0, 32 * gm["spacing"], (256, 3))
    trilinear = O.OracleField.grid(g["grid"], gm["origin"], gm["spacing"]).eval(pos)
    assert bits_equal(f.eval(pos), trilinear)


def test_decorated_tricubic_grid_matches_oracle():
    """PeriodicMagneticField and MagneticFieldList over a tricubic grid; the interpolation is set
    after the decorators are made (their device program is rebuilt before each use)."""
    c = O.load_case("interp_grids")
    m = c["meta"]
    grid = capi.Field.grid3f(c["grid"], m["origin"], m["spacing"])
    ext, org = [7 * s for s in m["spacing"]], m["origin"]
    per = capi.Field.periodic(grid, ext, org, reflective=True)
    lst = capi.Field.field_list([grid, capi.Field.uniform([1 * O.nG, 0, 0])])
    grid.set_interpolation("TRICUBIC")
    og = O.OracleField.grid(c["grid"], m["origin"], m["spacing"], interp=O.GRID_TRICUBIC)
    assert bits_equal(per.eval(c["pos"]), O.OracleField.periodic(og, ext, org, reflective=True).eval(c["pos"]))
    want = O.OracleField.field_list([og, O.OracleField.uniform([1 * O.nG, 0, 0])]).eval(c["pos"])
    assert bits_equal(lst.eval(c["pos"]), want)


def test_python_grid_classes():
    """The reference's classes: Grid3f / Grid1f setInterpolationType reach the device, for grid,
    modulated and JF12 fields."""
    from crpropa3_amd import (NEAREST_NEIGHBOUR, TRICUBIC, TRILINEAR, Grid1f, Grid3f, JF12Field, MagneticFieldGrid,
                              ModulatedMagneticFieldGrid, Vector3d)
    c = O.load_case("interp_grids")
    m = c["meta"]
    pos = c["pos"][:96]
    g = Grid3f(Vector3d(*m["origin"]), *m["shape"], Vector3d(*m["spacing"]))
    g.array()[:] = c["grid"]
    field = MagneticFieldGrid(g)

    def sample(f):
        return np.array([[v.x, v.y, v.z] for v in (f.getField(Vector3d(*p)) for p in pos)])

    g.setInterpolationType(TRICUBIC)
    assert g.getInterpolationType() == TRICUBIC
    assert bits_equal(sample(field), c["B_tricubic_periodic"][:96])
    g.setReflective(True)
    assert bits_equal(sample(field), c["B_tricubic_reflective"][:96])
    g.setInterpolationType(NEAREST_NEIGHBOUR)
    assert bits_equal(sample(field), c["B_nearest_reflective"][:96])
    mod = Grid1f(Vector3d(*m["mod_origin"]), *m["mod_shape"], Vector3d(*m["mod_spacing"]))
    mod.array()[:] = c["mod"]
    mod.setInterpolationType(TRICUBIC)
    g.setInterpolationType(TRICUBIC)
    mf = ModulatedMagneticFieldGrid(g, mod)  # Grid3f periodic, Grid1f reflective
    want = O.OracleField.modulated(c["grid"], m["origin"], m["spacing"], c["mod"], m["mod_origin"], m["mod_spacing"],
                                   interp=O.GRID_TRICUBIC, minterp=O.GRID_TRICUBIC).eval(pos)
    assert bits_equal(sample(mf), want)
    # JF12Field's turbulent part (JF12Field.cpp:340) through the same Grid3f, tricubic
    j = JF12Field()
    j.setTurbulentGrid(g)
    j.setUseTurbulentField(True)
    q = O.jf12_defaults()
    q.use_turbulent = 1
    og = O.OracleField.grid(c["grid"], m["origin"], m["spacing"], interp=O.GRID_TRICUBIC)
    og.s.type = O.FIELD_JF12
    og._keep.append(q)
    og.s.jf12 = O.C.cast(O.C.pointer(q), O.C.c_void_p)
    assert bits_equal(sample(j), og.eval(pos, libm=O.LIBM_CR))


@pytest.mark.parametrize("interp", ["TRICUBIC", "NEAREST_NEIGHBOUR"])
def test_jf12_turbulent_grid_interpolation_matches_oracle(interp):
    """JF12Field's turbulent part reads its grid with Grid::interpolate (src/magneticField/JF12Field.cpp:340):
    tricubic and nearest neighbour on the device (its own JF12 field type), bit for bit against the oracle
    with the correctly rounded functions, in field samples and a Cash-Karp run through the full model."""
    c = O.load_case("jf12")
    m = c["meta"]
    p, q = capi.jf12_default_params(), O.jf12_defaults()
    for x in (p, q):
        x.use_striated = 1
        x.use_turbulent = 1
    striated = O.jf12_striated(m["striated_seed"])
    dev = capi.Field.jf12(p, striated=striated, striated_spacing=0.1 * O.kpc, turbulent=c["grid"],
                          turbulent_spacing=m["turb_spacing"])
    dev.set_interpolation(interp)
    assert dev.interpolation() == interp
    orc = O.OracleField.jf12(q, striated, 0.1 * O.kpc, c["grid"], m["turb_spacing"])
    orc.s.interp = O.GRID_INTERP[interp]
    B = dev.eval(c["pos"])
    assert bits_equal(B, orc.eval(c["pos"], libm=O.LIBM_CR))
    trilinear = O.OracleField.jf12(q, striated, 0.1 * O.kpc, c["grid"], m["turb_spacing"]).eval(c["pos"],
                                                                                           libm=O.LIBM_CR)
    assert not np.array_equal(B, trilinear)
    t = O.load_case("ck_jf12")
    tm = t["meta"]
    n = len(t["energy"])
    prop = capi.Propagator("CK", dev, tm["tolerance"], tm["minStep"], tm["maxStep"],
                           max_length=tm["maxTrajectoryLength"])
    b = capi.Batch(n)
    b.upload(t["init_pos"][:n], O.normalise(t["init_dir"][:n]), t["energy"][:n], t["charge"][:n])
    prop.propagate(b, max_steps=40)
    st = O.OracleState(t["init_pos"][:n], t["init_dir"][:n], t["energy"][:n], t["charge"][:n])
    O.propagate(orc, O.PROP_CK, st, 40, tolerance=tm["tolerance"], min_step=tm["minStep"], max_step=tm["maxStep"],
                max_len=tm["maxTrajectoryLength"], libm=O.LIBM_CR)
    out = b.download()
    for key in ("pos", "dir", "next_step", "trajectory_length"):
        assert bits_equal(out[key], getattr(st, key)), key
