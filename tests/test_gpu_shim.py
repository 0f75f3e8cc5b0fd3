"""The cuOzBLAS C++ templates (include/cuozblas_mi355x.hpp) on the GPU: tests/shim_check.cpp calls
each binary64 and binary32 template and the C entry point it forwards to on the same inputs and
compares the outputs byte for byte (binary64, binary32 and the mixed routines)."""
import os
import subprocess

import pytest

from gpu_util import say, torch_cuda

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_shim_matches_c_entry_points():
    torch_cuda()
    exe = os.path.join(ROOT, "ozblas_amd", "_build", "shim_check")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    say("shim_check")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    say(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "shim check: all ok" in r.stdout
    assert r.stdout.count(" ok ") == 30
