"""CPU checks of the C-ABI library: it loads without a GPU and exports every
entry point declared in include/cubql_amd/cubql_capi.h; layout constants
match the reference's BinaryBVH<float,3> (bvh.h:71-111)."""
import ctypes as C
import os

import cubql_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    declared = cubql_amd._declared_symbols()
    assert len(declared) >= 10
    missing = declared - cubql_amd.exported_symbols()
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_abi_struct_layouts():
    assert C.sizeof(cubql_amd.BVH3f) == 32
    assert cubql_amd.BVH3f.primIDs.offset == 16
    assert cubql_amd.BVH3f.numPrims.offset == 24
    assert C.sizeof(cubql_amd.BuildConfig) == 12
    assert C.sizeof(cubql_amd.Box3f) == 24


def test_error_string_and_version():
    L = cubql_amd.lib()
    assert L.cubql_version().startswith(b"cubql_amd")
    # an invalid call is rejected without touching a device
    st = L.cubql_gpu_builder_f3(None, None, 0, None, None, None)
    assert st == -1
    assert b"null" in L.cubql_last_error_string()


def test_cpp_api_compiles_and_reports_errors():
    """The C++ host API (cubql.hpp) builds into a program that links the
    library; argument errors surface as std::runtime_error (no GPU work)."""
    import subprocess
    exe = os.path.join(ROOT, "cubql_amd", "cubql_cpp_check")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe, "abi"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK abi" in r.stdout
