"""The reference-side binding (crpropa3_amd/plugin): GpuPropagationBP / GpuPropagationCK modules of the
reference's Module interface that hand candidates to the C ABI.  Compiled here against the reference's
headers and linked with its library (a compile and link check: the reference does not travel to a GPU
machine), then run up to the device boundary, where a machine without a GPU reports the C ABI's error as
the reference's std::runtime_error."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
PLUGIN = os.path.join(ROOT, "crpropa3_amd", "plugin")

pytestmark = pytest.mark.skipif(
    not (os.path.isdir(os.path.join(REF, "include")) and
         os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libcrpropa_ref.a"))),
    reason="needs the reference's headers and its build (make -C oracle ref)")


def test_plugin_builds_and_runs_to_the_device_boundary():
    subprocess.check_call(["make", "-s", "-C", PLUGIN, "REF=" + REF])
    so = os.path.join(PLUGIN, "_build", "libcrpropa3_amd_plugin.so")
    syms = subprocess.check_output(["nm", "-DC", "--defined-only", so]).decode()
    for s in ("crpropa_amd::GpuPropagation::process(crpropa::Candidate*) const",
              "crpropa_amd::GpuPropagation::run(", "crpropa_amd::GpuPropagationCK::GpuPropagationCK(",
              "crpropa_amd::GpuPropagationBP::GpuPropagationBP(", "crpropa_amd::makeDeviceField("):
        assert s in syms, s
    env = dict(os.environ, KISS_LOG_LEVEL="0")
    out = subprocess.check_output([os.path.join(PLUGIN, "_build", "plugin_check")], env=env,
                                  stderr=subprocess.STDOUT, timeout=120).decode()
    assert "plugin_check: ran on the device" in out or "plugin_check: device unavailable" in out, out
