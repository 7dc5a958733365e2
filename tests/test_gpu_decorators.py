"""Field decorators on the device: PeriodicMagneticField (periodic and reflective), MagneticFieldList,
MagneticFieldEvolution and nestings of them (src/magneticField/MagneticField.cpp:41-79), through the C ABI.

* Field samples (tests/golden/decorators32, written by the reference): bit for bit against the
  reference where no transcendental function is involved, and against the oracle with the correctly
  rounded cos/pow the kernels use everywhere.
* Trajectories: Boris in a list (grid + uniform) and in an evolution with per-particle redshifts,
  bit for bit against the reference; adaptive Cash-Karp in a reflective periodic field bit for bit
  against the oracle (correctly rounded pow) and the reference's early snapshots.
* The mixed path and DiffusionSDE run through the same decorator program.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def device_grid():
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    return capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"]), gm


def device_fields(layout="cells"):
    c = O.load_case("decorators32")
    m = c["meta"]
    grid, gm = device_grid()
    grid.set_layout(layout)
    uni = capi.Field.uniform(m["uniform"])
    t = O.planewave_tables(O.load_case("planewave_simple2")["meta"])
    pw = capi.Field.planewave(t["k"], t["Ak"], t["xi"], t["kappa"], t["beta"])
    inner = capi.Field.field_list([grid, uni])
    return c, {
        "B_periodic": capi.Field.periodic(grid, m["extends"], m["origin"]),
        "B_reflective": capi.Field.periodic(grid, m["extends"], m["origin"], reflective=True),
        "B_list": capi.Field.field_list([grid, uni, pw]),
        "B_evolution": capi.Field.evolution(grid, m["evolution_m"]),
        "B_nested": capi.Field.field_list([capi.Field.periodic(grid, m["extends"], m["origin"], reflective=True),
                                           capi.Field.evolution(grid, m["nested_evolution_m"]),
                                           capi.Field.periodic(inner, m["nested_extends"])]),
        "B_evolution_list": capi.Field.evolution(inner, m["evolution_list_m"]),
    }


# samples whose arithmetic has no transcendental function: equal to the reference itself
REF_EXACT = {"B_periodic", "B_reflective", "B_nested"}


@pytest.mark.parametrize("layout", ["cells", "points"])
def test_decorator_field_samples(layout):
    c, dev = device_fields(layout)
    _, orc = O.decorator_fields()
    for name, f in dev.items():
        got = f.eval_z(c["pos"], c["z"])
        want = orc[name].eval(c["pos"], libm=O.LIBM_CR, z=c["z"])
        assert bits_equal(got, want), name
        rel = np.abs(got - c[name]).max() / np.abs(c[name]).max()
        print("%s (%s): bit-exact vs oracle(CR); max deviation from the reference %.1e" % (name, layout, rel))
        if name in REF_EXACT:
            assert bits_equal(got, c[name]), name
        else:
            assert rel < 1e-13, name


def test_decorator_one_argument_call():
    """getField(position) of a top-level evolution is the base class's zero; of the others, the same
    values as getField(position, z)."""
    c, dev = device_fields()
    assert bits_equal(dev["B_evolution"].eval(c["pos"]), c["B_evolution_noz"])
    for name in ("B_periodic", "B_reflective"):
        assert bits_equal(dev[name].eval(c["pos"]), c[name])


def run_golden(name, layout="cells", precision=capi.PRECISION_FP64):
    c = O.load_case(name)
    meta = c["meta"]
    grid, gm = device_grid()
    grid.set_layout(layout)
    if meta["field"].startswith("list"):
        f = capi.Field.field_list([grid, capi.Field.uniform([1 * O.nG, 0, 0])])
    elif meta["field"].startswith("periodic"):
        f = capi.Field.periodic(grid, meta["extends"], meta["origin"], reflective=True)
    else:
        f = capi.Field.evolution(grid, meta["evolution_m"])
    a = O.propagator_args(meta)
    kind = "BP" if a["prop_type"] == O.PROP_BP else "CK"
    prop = capi.Propagator(kind, f, a["tolerance"], a["min_step"], a["max_step"],
                           max_length=meta.get("maxTrajectoryLength", 0.0), precision=precision)
    b = capi.Batch(len(c["energy"]))
    b.upload(c["init_pos"], O.normalise(c["init_dir"]), c["energy"], c["charge"], redshift=c.get("redshift"))
    out, done = [], 0
    for s in c["snap_steps"]:
        prop.propagate(b, max_steps=int(s) - done)
        done = int(s)
        out.append(b.download())
    return c, out


@pytest.mark.parametrize("name", ["bp_fixed_list32", "bp_fixed_evolution32", "ck_reflective32"])
@pytest.mark.parametrize("layout", ["cells", "points"])
def test_decorator_trajectories(name, layout):
    c, out = run_golden(name, layout)
    import test_oracle as T
    _, snaps, _ = T.run_case(name, O.LIBM_CR)
    for si, g in enumerate(out):
        for key in ("pos", "dir", "current_step", "next_step", "trajectory_length", "time"):
            assert bits_equal(g[key], snaps[si][key]), (name, si, key)
        ref_exact = name != "ck_reflective32" or si < 2
        if ref_exact:  # the adaptive case leaves the reference after glibc's first pow rounding
            assert bits_equal(g["pos"], c["pos"][si]) and bits_equal(g["dir"], c["dir"][si]), (name, si)


def test_decorator_mixed_precision_short_horizon():
    c, ref = run_golden("bp_fixed_list32")
    _, mix = run_golden("bp_fixed_list32", precision=capi.PRECISION_MIXED)
    e = np.max(np.abs(mix[1]["pos"] - ref[1]["pos"]), axis=1) / np.max(np.abs(ref[1]["pos"]), axis=1)
    print("mixed list field, 10 steps: median relative position error %.2e" % np.median(e))
    assert np.median(e) < 1e-6
    c, ref = run_golden("bp_fixed_evolution32")
    _, mix = run_golden("bp_fixed_evolution32", precision=capi.PRECISION_MIXED)
    e = np.max(np.abs(mix[1]["pos"] - ref[1]["pos"]), axis=1) / np.max(np.abs(ref[1]["pos"]), axis=1)
    assert np.median(e) < 1e-6


def test_decorator_sde_matches_oracle():
    """DiffusionSDE in an evolving periodic field: the device's Philox stream against the oracle."""
    grid, gm = device_grid()
    L = gm["n"] * gm["spacing"]
    f = capi.Field.evolution(capi.Field.periodic(grid, [1.3 * L, 0.8 * L, L], [0, 0, 0], reflective=True), 0.7)
    n = 256
    rng = np.random.default_rng(8)
    pos = np.tile([L / 2] * 3, (n, 1))
    d = O.normalise(rng.normal(size=(n, 3)))
    E = 10 ** rng.uniform(16, 18, n) * O.eV
    q = np.full(n, O.eplus)
    z = rng.uniform(0, 2, n)
    prop = capi.Propagator("SDE", f, 1e-4, 10 * O.pc, 1 * O.kpc, epsilon=0.1, alpha=1 / 3., scale=1.0, seed=5)
    b = capi.Batch(n)
    b.upload(pos, d, E, q, redshift=z)
    prop.propagate(b, max_steps=50)
    g = b.download()
    og = O.OracleField.grid(O.load_case("grid32")["grid"], gm["origin"], gm["spacing"])
    of = O.OracleField.evolution(O.OracleField.periodic(og, [1.3 * L, 0.8 * L, L], [0, 0, 0], reflective=True), 0.7)
    st = O.OracleState(pos, d, E, q, normalised=True)
    st.redshift[:] = z
    steps = np.zeros(n, dtype=np.uint32)
    O.sde_propagate(of, st, 50, 1e-4, 10 * O.pc, 1 * O.kpc, epsilon=0.1, alpha=1 / 3., scale=1.0, seed=5,
                    rng=O.SDE_PHILOX, sde_step=steps)
    assert bits_equal(g["pos"], st.pos)
    assert bits_equal(g["dir"], st.dir)


def test_decorator_limits_rejected():
    grid, _ = device_grid()
    lst = capi.Field.field_list([grid] * 8)
    with pytest.raises(capi.CrpError, match="at most 8"):
        capi.check(capi.lib().crp_field_list_add(lst.handle, grid.handle))
    f = grid
    for _ in range(3):
        f = capi.Field.periodic(f, [1e22, 1e22, 1e22])
    with pytest.raises(capi.CrpError, match="nested"):
        capi.Field.periodic(f, [1e22, 1e22, 1e22])
