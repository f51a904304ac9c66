"""Runs the drop-in C++ API program (tests/cpp/dropin_test.cc, built by
__graft_entry__.build() with g++ against libqphix_amd.so) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_dropin_cpp_api():
    exe = os.path.join(HERE, "cpp", "dropin_test")
    assert os.path.exists(exe), "build the drop-in test with __graft_entry__.build()"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=110)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "DROPIN OK" in p.stdout
