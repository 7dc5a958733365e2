"""The reference's public single-step functions on the device (include/crpropa/module/PropagationBP.h:89-116,
PropagationCK.h:74-91): PropagationBP::dY, errorEstimation, tryStep, getFieldAtPosition and PropagationCK::dYdt,
tryStep, getFieldAtPosition, through the C ABI and the host classes, bit for bit against the oracle's
restatement (src/module/PropagationBP.cpp:8-39,152-159, src/module/PropagationCK.cpp:28-61), and
consistent with the propagation kernels."""
import numpy as np
import pytest

import oracle_ctypes as O
import crpropa3_amd as crp
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float64).view(np.uint64),
                          np.ascontiguousarray(b, np.float64).view(np.uint64))


def fields():
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    dg = capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"])
    og = O.OracleField.grid(g["grid"], gm["origin"], gm["spacing"])
    t = O.planewave_tables(O.load_case("planewave_c3")["meta"])
    dp = capi.Field.planewave(t["k"], t["Ak"], t["xi"], t["kappa"], t["beta"])
    op = O.OracleField.planewave(t)
    L = gm["n"] * gm["spacing"]
    de = capi.Field.evolution(capi.Field.periodic(dg, [1.5 * L, L, 0.7 * L], reflective=True), 1.2)
    oe = O.OracleField.evolution(O.OracleField.periodic(og, [1.5 * L, L, 0.7 * L], reflective=True), 1.2)
    return {"grid32": (dg, og, L), "planewave_c3": (dp, op, 200 * O.pc), "evolution_periodic": (de, oe, L)}


@pytest.mark.parametrize("name", ["grid32", "planewave_c3", "evolution_periodic"])
def test_single_step_functions_match_oracle(name):
    dev, orc, L = fields()[name]
    rng = np.random.default_rng(21)
    n = 4096
    x = rng.uniform(-L, 2 * L, size=(n, 3))
    u = O.normalise(rng.normal(size=(n, 3)))
    E = 10 ** rng.uniform(17, 20.5, n) * O.eV
    q = rng.choice([1.0, 2.0, 26.0, -1.0], n) * O.eplus
    z = rng.uniform(0, 2, n)
    step = rng.uniform(0.01, 1.0, n) * L / 10
    m = E / (O.c_light * O.c_light)
    bp = capi.Propagator("BP", dev, 1e-4, 1 * O.kpc, 1 * O.Mpc)
    ck = capi.Propagator("CK", dev, 1e-4, 1 * O.kpc, 1 * O.Mpc)
    got = bp.step("dY", x, u, step, z, q, m)
    want = O.step_batch(orc, O.STEP_BP_DY, x, u, step, z, q, m)
    assert bits_equal(got[0], want[0]) and bits_equal(got[1], want[1])
    got = bp.step("tryStep", x, u, step, z, q, m)
    want = O.step_batch(orc, O.STEP_BP_TRY, x, u, step, z, q, m)
    for a, b in zip(got, want):
        assert bits_equal(a, b)
    assert np.all(got[2][:, 0] == got[2][:, 1]) and np.all(got[3] == got[2])  # Y(errorEstimation)
    got = ck.step("dYdt", x, u, 0, z, q, E)
    want = O.step_batch(orc, O.STEP_CK_DYDT, x, u, 0, z, q, E)
    assert bits_equal(got[0], want[0]) and bits_equal(got[1], want[1])
    t = step / O.c_light
    got = ck.step("tryStep", x, u, t, z, q, E)
    want = O.step_batch(orc, O.STEP_CK_TRY, x, u, t, z, q, E)
    for a, b in zip(got, want):
        assert bits_equal(a, b)


def test_dy_is_the_fixed_step_process():
    """One fixed-step process() call = dY, then the two direction normalisations
    (src/module/PropagationBP.cpp:121, src/ParticleState.cpp:30-32)."""
    dev, _, L = fields()["grid32"]
    rng = np.random.default_rng(4)
    n = 1024
    x = rng.uniform(0, L, size=(n, 3))
    u = O.normalise(rng.normal(size=(n, 3)))
    E = 10 ** rng.uniform(18, 20, n) * O.eV
    q = np.full(n, O.eplus)
    h = 0.05 * O.Mpc
    bp = capi.Propagator("BP", dev, 0.42, h, h)
    ox, ou = bp.step("dY", x, u, h, 0.0, q, E / (O.c_light * O.c_light))
    b = capi.Batch(n)
    b.upload(x, u, E, q)
    bp.propagate(b, max_steps=1)
    g = b.download()
    assert bits_equal(g["pos"], ox)
    assert bits_equal(g["dir"], O.normalise(O.normalise(ou)))


def test_host_class_single_step_api():
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    grid = crp.Grid3f(crp.Vector3d(0.0), gm["n"], gm["spacing"])
    np.asarray(grid.array())[...] = g["grid"]
    field = crp.MagneticFieldGrid(grid)
    og = O.OracleField.grid(g["grid"], gm["origin"], gm["spacing"])
    pos = crp.Vector3d(0.3 * crp.Mpc, 0.7 * crp.Mpc, 1.1 * crp.Mpc)
    d = crp.Vector3d(0.48, -0.6, 0.64)
    E, q = 3 * crp.EeV, 2 * crp.eplus
    bp = crp.PropagationBP(field, 1e-4, 1 * crp.kpc, 1 * crp.Mpc)
    ck = crp.PropagationCK(field, 1e-4, 1 * crp.kpc, 1 * crp.Mpc)
    # getFieldAtPosition (test/testPythonExtension.py:90-91 with a device field)
    fa = field.getField(pos, 0.0)
    assert fa == bp.getFieldAtPosition(pos, 0.0) == ck.getFieldAtPosition(pos, 0.0)
    want = og.eval([[pos.x, pos.y, pos.z]])[0]
    assert (fa.x, fa.y, fa.z) == tuple(want)
    x = np.array([[pos.x, pos.y, pos.z]])
    u = np.array([[d.x, d.y, d.z]])
    m = E / (crp.c_light * crp.c_light)
    y = bp.dY(pos, d, 0.1 * crp.Mpc, 0.0, q, m)
    w = O.step_batch(og, O.STEP_BP_DY, x, u, 0.1 * crp.Mpc, 0.0, q, m)
    assert (y.x.x, y.x.y, y.x.z, y.u.x, y.u.y, y.u.z) == tuple(w[0][0]) + tuple(w[1][0])
    ps = crp.ParticleState(crp.nucleusId(4, 2), E, pos, d)
    out, err = crp.PropagationBP.Y(), crp.PropagationBP.Y()
    bp.tryStep(crp.PropagationBP.Y(pos, d), out, err, 0.1 * crp.Mpc, ps, 0.0, q, m)
    w = O.step_batch(og, O.STEP_BP_TRY, x, u, 0.1 * crp.Mpc, 0.0, q, m)
    assert (out.x.x, out.u.z, err.x.x, err.u.y) == (w[0][0][0], w[1][0][2], w[2][0][0], w[3][0][1])
    S = bp.errorEstimation(out.x, crp.Vector3d(w[0][0][0] + 1.0, w[0][0][1], w[0][0][2]), 0.1 * crp.Mpc)
    assert S == pytest.approx(1.0 / (0.1 * crp.Mpc * 0.75), rel=1e-12)
    dy = ck.dYdt(crp.PropagationCK.Y(pos, d), ps, 0.0)
    w = O.step_batch(og, O.STEP_CK_DYDT, x, u, 0.0, 0.0, ps.getCharge(), E)
    assert (dy.x.x, dy.u.x, dy.u.y, dy.u.z) == (w[0][0][0], w[1][0][0], w[1][0][1], w[1][0][2])
    ck.tryStep(crp.PropagationCK.Y(pos, d), out, err, 1e3 * crp.kpc / crp.c_light, ps, 0.0)
    w = O.step_batch(og, O.STEP_CK_TRY, x, u, 1e3 * crp.kpc / crp.c_light, 0.0, ps.getCharge(), E)
    assert (out.x.y, out.u.z, err.x.x, err.u.z) == (w[0][0][1], w[1][0][2], w[2][0][0], w[3][0][2])
