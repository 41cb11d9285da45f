"""GPU run of the C++ host API (include/cubql_amd/cubql.hpp): every builder
through cuBQL::gpuBuilder / cuBQL::cuda::*, with the default resource and a
user GpuMemoryResource, closest-point results equal to a float brute force,
and every byte allocated through the user resource returned by free()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_api_on_gpu():
    print("[gpu] C++ API program", flush=True)
    exe = os.path.join(ROOT, "cubql_amd", "cubql_cpp_check")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK gpu" in r.stdout
