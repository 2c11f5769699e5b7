"""The C-ABI libraries load and export every function their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from gicops_amd import _build, _ffi

INCLUDE = Path(__file__).resolve().parent.parent / "include"
DECL = re.compile(r"^\s*(?:int|void|const\s+char\s*\*|char\s*\*)\s+\**\s*(gx_\w+)\s*\(", re.M)


def declared(header: str) -> list[str]:
    return sorted(set(DECL.findall((INCLUDE / header).read_text())))


def _hip_lib_path():
    p = _build.LIBDIR / "libgicops_hip.so"
    if not p.exists():
        _build.build_hip()
    return p


def test_headers_declare_functions():
    assert len(declared("gicops_amd.h")) >= 10
    assert len(declared("gicops_index.h")) >= 8


def test_host_library_exports_index_abi():
    lib = _ffi.host_lib()
    for name in declared("gicops_index.h"):
        assert hasattr(lib, name), name


def test_hip_library_exports_search_abi():
    lib = C.CDLL(str(_hip_lib_path()))
    for name in declared("gicops_amd.h"):
        assert hasattr(lib, name), name


def test_status_codes_match_reference_values():
    # slmerr.h:24-46: SLM_ERROR_BASE = -3000
    assert _ffi.GX_ERR_INVLD_PARAM == -3002
    assert _ffi.GX_ERR_INVLD_INDEX == -3003
    assert _ffi.GX_ERR_INVLD_PTR == -3008


def test_host_abi_rejects_bad_arguments():
    lib = _ffi.host_lib()
    p = _ffi.gx_db_params()
    p.min_len, p.max_len, p.maxz, p.scale, p.min_mass, p.max_mass = 7, 6, 1, 100, 500, 5000
    h = C.c_void_p()
    assert lib.gx_db_create(C.byref(p), C.byref(h)) == _ffi.GX_ERR_INVLD_PARAM
    assert lib.gx_db_create(None, C.byref(h)) == _ffi.GX_ERR_INVLD_PTR
    b = _ffi.gx_bucket()
    assert lib.gx_db_get_bucket(None, 0, C.byref(b)) == _ffi.GX_ERR_INVLD_PTR


def test_gpu_entry_points_fail_cleanly_without_device():
    """Without a visible GPU, context creation returns a status, never aborts."""
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a GPU is visible")
    lib = _ffi.hip_lib()
    ctx = C.c_void_p()
    st = lib.gx_ctx_create(0, C.byref(ctx))
    assert st in (_ffi.GX_ERR_NO_DEVICE, _ffi.GX_ERR_DEVICE)
    assert not ctx.value
