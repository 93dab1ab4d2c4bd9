"""The C ABI: libpxh.so loads and exports every entry point declared in
include/pxh.h (no compute call is made here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from plasmonx_amd import _lib


def declared():
    text = open(os.path.join(ROOT, "include", "pxh.h")).read()
    return sorted(set(re.findall(r"\bint\s+(pxh_\w+)\s*\(", text)))


def test_header_declares_the_binding_set():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpxh.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_fortran_interface_binds_every_symbol():
    text = open(os.path.join(ROOT, "plasmonx_amd", "fortran", "pxh.f90")).read()
    for name in declared():
        assert f'name="{name}"' in text, name


def test_null_context_is_rejected_without_device():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpxh.so not built")
    lib = _lib.load()
    assert lib.pxh_solve_spectrum(None, 1, None, 0.0, 1, 7, 0, 0.0, 0, 0, None, None, None, None, None) == \
        _lib.PXH_EINVAL
    buf = ctypes.create_string_buffer(64)
    assert lib.pxh_last_error(None, buf, 64) == _lib.PXH_OK
