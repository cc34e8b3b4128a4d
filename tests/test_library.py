"""The C-ABI library: it loads and exports every symbol that
include/ephemeris_hip.h declares, and the Python binding declares exactly
those.  No GPU call is made."""
import ctypes as C
import re

from ephemeris_amd import _lib
from tests.util import ROOT


def declared_symbols():
    text = (ROOT / "include" / "ephemeris_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eph_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_abi_version_and_error_string():
    L = _lib.lib()
    assert L.eph_abi_version() == 1
    assert isinstance(L.eph_last_error(), bytes)


def test_null_handles_are_rejected():
    L = _lib.lib()
    assert L.eph_accel(None, 0) == _lib.EPH_EINVAL
    assert L.eph_integrate(None, 300.0, 1) == _lib.EPH_EINVAL
    assert L.eph_system_destroy(None) == _lib.EPH_EINVAL
    assert b"null" in L.eph_last_error()
