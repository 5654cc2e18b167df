"""liblm.so loads without a GPU and exports every entry point include/lm.h declares."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lm.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(lm_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def liblm():
    from lagrangian_microbes_amd import _lib, build
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    return _lib.load()


def test_header_declares_entry_points():
    names = declared()
    for must in ("lm_advect_rk4_f64", "lm_query_pairs", "lm_rps_update", "lm_field_blend_f64",
                 "lm_philox_uniforms", "lm_last_error"):
        assert must in names


def test_every_declared_symbol_exported(liblm):
    raw = ctypes.CDLL(liblm._name)
    for name in declared():
        assert hasattr(raw, name), f"{name} declared in lm.h but not exported"


def test_binding_covers_header(liblm):
    from lagrangian_microbes_amd import _lib
    assert sorted(_lib.EXPORTED) == declared()


def test_abi_version_and_argument_errors(liblm):
    assert liblm.lm_abi_version() == 6
    # argument validation runs on the host before any launch: no GPU is touched
    rc = liblm.lm_advect_rk4_f64(-1, None, None, None, 0, None, 0, 1.0, None, None, 0, 0.0, 0, 0, 0, None, None)
    assert rc != 0 and b"negative" in liblm.lm_last_error()
    rc = liblm.lm_philox_uniforms(4, 0, 0, 0, 0, 0, 0, 7, None, None)
    assert rc != 0 and b"word" in liblm.lm_last_error()
    rc = liblm.lm_query_pairs(None, 0, None, None, 0.1, 2.0, None, None, 0, None, None)
    assert rc != 0 and b"workspace" in liblm.lm_last_error()
    rc = liblm.lm_rps_ws_reserve(None, 1000, None)
    assert rc != 0 and b"workspace" in liblm.lm_last_error()


def test_adjoint_argument_errors(liblm):
    from lagrangian_microbes_amd._lib import LmGrid
    import numpy as np
    g = LmGrid(5, 5, 0.0, 0.0, 4.0, 4.0)
    lev = np.array([0, 0, 1, 1, 2], dtype=np.int32)  # 2 steps; level 2 + 1 does not exist with 3 levels
    rc = liblm.lm_adjoint_forward(10, None, None, None, 3, ctypes.byref(g), None, None,
                                  lev.ctypes.data_as(ctypes.c_void_p), 2, 3600.0, 16, None, None, 0, None, None)
    assert rc != 0 and b"level 2" in liblm.lm_last_error()
    rc = liblm.lm_adjoint_backward(10, None, None, 3, ctypes.byref(g), None, None,
                                   lev.ctypes.data_as(ctypes.c_void_p), 2, 3600.0, 0, None, 0, None, None, None,
                                   None)
    assert rc != 0 and b"checkpoint" in liblm.lm_last_error()
    ok = np.array([0, 0, 1, 1, 1], dtype=np.int32)
    rc = liblm.lm_adjoint_backward(10, None, None, 3, ctypes.byref(g), None, None,
                                   ok.ctypes.data_as(ctypes.c_void_p), 2, 3600.0, 16, None, 0, None, None, None,
                                   None)
    assert rc != 0 and b"NULL" in liblm.lm_last_error()
    # scratch: 16 start positions of 16 bytes per particle, one 16-byte box per 256-particle block,
    # 33 float64 stage fields of 5x5 (u, v) nodes (one for a steady field) and a slab of 16 floats per
    # level pair (2 of 3 levels; 1 when steady) and 4x4 cell
    base = 16 * 1000 * 16 + 4 * 16
    assert liblm.lm_adjoint_scratch_bytes(1000, 16, ctypes.byref(g), 3) == base + 33 * 25 * 16 + 2 * 16 * 64
    assert liblm.lm_adjoint_scratch_bytes(1000, 16, ctypes.byref(g), 1) == base + 25 * 16 + 1 * 16 * 64
    assert liblm.lm_adjoint_scratch_bytes(-1, 16, ctypes.byref(g), 3) == -1
