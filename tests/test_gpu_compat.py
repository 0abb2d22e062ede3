"""The reference-name layer (include/mpres_compat.hpp) run on the GPU: tests/host/compat_run.cpp
calls every cuda::mp_* name the way the reference's harnesses do, on seeded inputs, and
compares each result bit for bit with the C-ABI call it must forward to."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_compat_layer_runs(tmp_path):
    exe = tmp_path / "compat_run"
    lib = os.path.join(ROOT, "mpres-blas_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "host", "compat_run.cpp"), "-o", str(exe),
           "-L", lib, "-lmpres_hip", "-L", "/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("[case] compat layer run", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "compat run ok: 10 checks" in r.stdout
