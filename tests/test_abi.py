"""CPU checks of the C ABI: the library loads and exports every symbol include/hesship.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hesship.h")
LIB = os.path.join(ROOT, "rstsr-showcase-hessian_amd", "hesship", "libhesship.so")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hh_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "hh_rijk_get_response_bra" in syms and len(syms) >= 8


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhesship.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhesship.so not built")
def test_binding_signatures_cover_header():
    from hesship import _lib
    assert set(declared_symbols()) <= set(_lib.SIGNATURES), set(declared_symbols()) - set(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhesship.so not built")
def test_version_and_invalid_argument_paths():
    """Host-side argument checks run without touching a GPU."""
    from hesship import _lib
    L = _lib.lib()
    assert b"gfx950" in L.hh_version()
    h = ctypes.c_void_p()
    rc = L.hh_rijk_create(None, 4, 3, 1.0, 1.0, 3, None, ctypes.byref(h))
    assert rc == _lib.HH_ERR_INVALID and b"nspin" in L.hh_last_error()
    rc = L.hh_rijk_get_response_bra(None, None, 1, None, 0)
    assert rc == _lib.HH_ERR_INVALID


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhesship.so not built")
def test_libxc_to_xcfun_map_host_entry():
    """The host entry point reproduces the reference's table (xc_deriv.rs:17-40) and rejects
    derivative orders it does not cover."""
    from hesship import _lib
    from hesship.numint import libxc_to_xcfun_mapping
    from tests.test_oracle_uks_xceff import REF_MAP
    for den, ref in REF_MAP.items():
        assert list(libxc_to_xcfun_mapping(den, 2)) == ref
    assert list(libxc_to_xcfun_mapping("rho", 2)) == [0, 1, 2, 3, 4, 5]
    n = ctypes.c_int64()
    assert _lib.lib().hh_xceff_libxc_to_xcfun_map(1, 3, None, 0, ctypes.byref(n)) == _lib.HH_ERR_INVALID
    buf = (ctypes.c_int64 * 2)()
    assert _lib.lib().hh_xceff_libxc_to_xcfun_map(1, 2, buf, 2, ctypes.byref(n)) == _lib.HH_ERR_INVALID
