"""The C-ABI library loads and exports every function include/yk.h declares.

No compute calls: this runs without a GPU.
"""
import ctypes
import re
from pathlib import Path

import pytest

from yukawamd_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "yk.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(yk_[a-z_0-9]+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("yk_create", "yk_destroy", "yk_accel", "yk_newton", "yk_run", "yk_put_state",
                 "yk_get_state", "yk_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} is not built (python -m yukawamd_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_fortran_module_binds_the_header():
    """yk_bind.f90 (the Fortran host's interface module) binds every C entry."""
    f90 = (HEADER.parent.parent / "yukawamd_amd" / "fortran" / "yk_bind.f90").read_text()
    bound = set(re.findall(r"bind\(C,\s*name='(yk_[a-z_0-9]+)'\)", f90))
    assert set(declared_functions()) <= bound, set(declared_functions()) - bound


def test_abi_version_and_error_string():
    lib = _lib.load()
    assert lib.yk_abi_version() == 1
    assert isinstance(lib.yk_last_error(), bytes)


def test_invalid_params_rejected_without_device():
    lib = _lib.load()
    from yukawamd_amd.params import SimParams
    p = _lib.params_struct(SimParams(sim_type="pure-ion", xl=(10.0, 10.0, -1.0), rccut=3.0,
                                     zi=8.0, ai=16.0))
    h = ctypes.c_void_p()
    rc = lib.yk_create(ctypes.byref(p), 4, None, None, ctypes.byref(h))
    assert rc == 1 and b"xl(3)" in lib.yk_last_error()
