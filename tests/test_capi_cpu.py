"""The C ABI library loads without a GPU, exports every function that
include/*.h declares, and validates arguments like the reference's setters."""
import ctypes as C
import glob
import os
import re

import pytest

from crpropa3_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\*?\s*(crp_[a-z0-9_]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version():
    assert capi.lib().crp_abi_version() == 12


@pytest.mark.parametrize("kind", ["BP", "CK"])
def test_parameter_validation(kind):
    """test/testPropagation.cpp:61-81 / :295-315: the same cases raise."""
    with pytest.raises(capi.CrpError, match="minStep > maxStep"):
        capi.Propagator(kind, None, 0.42, 10, 0)
    with pytest.raises(capi.CrpError, match="target error not in range 0-1"):
        capi.Propagator(kind, None, 42.0, 10, 20)
    with pytest.raises(capi.CrpError, match="minStep < 0"):
        capi.Propagator(kind, None, 0.42, -1.0, 20)


def test_planewave_needs_two_modes():
    with pytest.raises(capi.CrpError, match="Nm <= 1"):
        capi.Field.planewave([1.0], [1.0], [[1, 0, 0]], [[0, 0, 1]], [0.0])


def test_last_error_is_cleared_per_message():
    with pytest.raises(capi.CrpError) as e:
        capi.Propagator("BP", None, 2.0, 0, 1)
    assert "PropagationBP" in str(e.value)
