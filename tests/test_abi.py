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
    assert liblm.lm_abi_version() == 3
    # argument validation runs on the host before any launch: no GPU is touched
    rc = liblm.lm_advect_rk4_f64(-1, None, None, None, 0, None, 0, 1.0, None, None, 0, 0.0, 0, 0, 0, None, None)
    assert rc != 0 and b"negative" in liblm.lm_last_error()
    rc = liblm.lm_philox_uniforms(4, 0, 0, 0, 0, 0, 0, 7, None, None)
    assert rc != 0 and b"word" in liblm.lm_last_error()
    rc = liblm.lm_query_pairs(None, 0, None, None, 0.1, 2.0, None, None, 0, None, None)
    assert rc != 0 and b"workspace" in liblm.lm_last_error()
