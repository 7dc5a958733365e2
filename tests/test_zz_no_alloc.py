"""crp_field_eval_noalloc (field evaluation staged through the library's fixed per-device buffer, no
device allocation) against crp_field_eval_ex on the same fields and positions: bit for bit, on two
calls in a row with the second smaller, for the FP64 and the mixed evaluation."""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu

MAX = 131072  # CRP_EVAL_NOALLOC_MAX of include/crpropa3_amd.h


def _eval(fn, field, pos, precision):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    B = np.full_like(pos, np.nan)
    dp = C.POINTER(C.c_double)
    capi.check(fn(field.handle, pos.ctypes.data_as(dp), pos.shape[0], B.ctypes.data_as(dp), precision))
    return B


def fields():
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    grid = capi.Field.grid3f(g["grid"], gm["origin"], gm["spacing"])
    c = O.load_case("planewave_c3")
    t = O.planewave_tables(c["meta"])
    pw = capi.Field.planewave(t["k"], t["Ak"], t["xi"], t["kappa"], t["beta"])
    return {"grid": (grid, gm["n"] * gm["spacing"]), "planewave": (pw, 100 * O.pc)}


@pytest.mark.parametrize("precision", [capi.PRECISION_FP64, capi.PRECISION_MIXED])
def test_noalloc_eval_equals_eval_ex_two_calls(precision):
    L = capi.lib()
    rng = np.random.default_rng(17)
    for name, (f, size) in fields().items():
        for n in (MAX, 1000):  # the full capacity, then a smaller call over the same buffer
            pos = rng.uniform(-size, 2 * size, size=(n, 3))
            ref = _eval(L.crp_field_eval_ex, f, pos, precision)
            got = _eval(L.crp_field_eval_noalloc, f, pos, precision)
            print("%s precision %d n %d: max |B| %.3e" % (name, precision, n, np.abs(ref).max()), flush=True)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (name, n)


def test_noalloc_eval_rejects_oversize_and_empty():
    L = capi.lib()
    f = capi.Field.uniform([0, 0, 1e-13])
    with pytest.raises(capi.CrpError, match="staging capacity"):
        _eval(L.crp_field_eval_noalloc, f, np.zeros((MAX + 1, 3)), capi.PRECISION_FP64)
    assert _eval(L.crp_field_eval_noalloc, f, np.zeros((0, 3)), capi.PRECISION_FP64).shape == (0, 3)
    B = _eval(L.crp_field_eval_noalloc, f, np.zeros((3, 3)), capi.PRECISION_FP64)
    assert np.array_equal(B, np.tile([0, 0, 1e-13], (3, 1)))
