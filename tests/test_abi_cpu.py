"""CPU checks of the C ABI: the library loads, exports every declared symbol, and
its structures and constants match the binding (no GPU call is made)."""
import ctypes
import os
import re

import pytest

from pypatools_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pypat_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding covers all of them
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_struct_layout_matches_binding():
    assert _lib.layout_ok()
    assert _lib.lib().pp_abi_version() == 3


def test_library_was_built_from_these_kernel_sources():
    """The sha256 compiled into the library equals that of the kernel sources on disk (a stale build
    would otherwise be timed and profiled under the sources' name)."""
    from pypatools_amd import build
    assert re.fullmatch(r"[0-9a-f]{64}", _lib.kernel_digest())
    assert _lib.kernel_digest() == build.kernel_source_digest()


def test_constants_match_reference_expressions():
    from oracle import kernels as K
    from oracle.constants import CLIGHT
    c = _lib.constants()
    assert c[0] == CLIGHT
    assert c[1] == CLIGHT ** 2
    assert c[2] == 0.9999 * CLIGHT
    assert c[3:6] == [K.YOSHIDA_C1, K.YOSHIDA_C2, K.YOSHIDA_C3]
    assert c[6:9] == [K.YOSHIDA_D1, K.YOSHIDA_D2, K.YOSHIDA_D3]


def test_argument_validation_happens_before_any_device_call():
    lib = _lib.lib()
    assert lib.pp_get_constants(None, 9) == _lib.PP_ERR_INVALID
    assert b"constants" in lib.pp_last_error()
    a, b = ctypes.c_int64(), ctypes.c_int64()
    assert lib.pp_struct_sizes(None, ctypes.byref(b)) == _lib.PP_ERR_INVALID


def test_header_defines_match_binding():
    text = open(os.path.join(ROOT, "include", "pypat_hip.h")).read()
    defs = dict(re.findall(r"#define\s+(PP_[A-Z_]+)\s+(\d+)", text))
    assert int(defs["PP_MOMENT_SUMS"]) == _lib.MOMENT_SUMS
    assert int(defs["PP_MOMENT_CENTRAL"]) == _lib.MOMENT_CENTRAL
