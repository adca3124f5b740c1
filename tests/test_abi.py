"""CPU tests of the C-ABI boundary: the library loads, exports every symbol
include/femhip.h declares, and fails cleanly (no crash) without a GPU."""
import ctypes as C
import os
import re

import numpy as np

from femhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "femhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(fem_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    names = declared_functions()
    bound = sorted(n for n, _, _ in _lib.SIGNATURES)
    assert names == bound


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_sizes_match_c_layout():
    # sizes implied by the header (x86-64 SysV layout)
    assert C.sizeof(_lib.fem_mesh) == 32
    assert C.sizeof(_lib.fem_assembly_params) == 48
    assert C.sizeof(_lib.fem_rd_params) == 48
    assert C.sizeof(_lib.fem_solver_opts) == 64
    assert C.sizeof(_lib.fem_step_stats) == 48
    assert C.sizeof(_lib.fem_ctx_info) == 24


def test_version_defaults_and_errors_without_touching_a_device():
    lib = _lib.load()
    assert lib.fem_abi_version() == 1
    o = _lib.fem_solver_opts()
    lib.fem_opts_default(C.byref(o))
    assert o.precision == _lib.FEM_PREC_FP32_IR and o.nrhs == 1 and o.tol == 1e-10
    # invalid arguments are rejected before any HIP call
    h = C.c_void_p()
    assert lib.fem_ctx_create(None, 0, C.byref(h)) == _lib.FEM_ERR_INVALID
    assert "invalid mesh" in lib.fem_last_error().decode()
    m = _lib.fem_mesh(3, 1, None, None, None)
    assert lib.fem_ctx_create(C.byref(m), 0, C.byref(h)) == _lib.FEM_ERR_INVALID
    xyz = np.zeros(9, np.float32)
    tri = np.array([0, 1, 2], np.uint32)
    onb = np.zeros(3, np.uint8)
    m = _lib.fem_mesh(3, 1, _lib.ptr(xyz, C.c_float), _lib.ptr(tri, C.c_uint32), _lib.ptr(onb, C.c_uint8))
    assert lib.fem_ctx_create(C.byref(m), 7, C.byref(h)) == _lib.FEM_ERR_INVALID
    assert lib.fem_solver_advance(None, 1, None) == _lib.FEM_ERR_INVALID
    assert lib.fem_ctx_info_get(None, None) == _lib.FEM_ERR_INVALID


def test_device_count_is_safe_without_gpu():
    lib = _lib.load()
    c = C.c_int32(-1)
    assert lib.fem_device_count(C.byref(c)) == 0
    assert c.value >= 0
