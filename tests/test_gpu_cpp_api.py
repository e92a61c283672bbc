"""Runs the C++ API check program (tests/cpp/test_turbo_quant_index.cpp,
built by vectorcache_amd.build) on the GPU: the reference's index tests
restated against vectorcache::TurboQuantIndex over the C ABI."""
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_turbo_quant_index():
    exe = ROOT / "tests" / "cpp" / "test_turbo_quant_index"
    assert exe.exists(), "build first: python -m vectorcache_amd.build"
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("OK")
