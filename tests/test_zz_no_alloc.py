"""The allocation-free entry point eph_accel_noalloc against eph_accel (the
resident-batch path: eph_system_create + eph_accel + eph_system_download),
through the C ABI: two calls in a row on the same device array, the second
with a smaller system, each bit-identical to the resident path; size limits
are rejected before anything runs.  The caller's device buffers for the
inputs and outputs come from hipMalloc of the HIP runtime the library itself
is linked against (a second HIP runtime in the process, such as the one a
PyTorch wheel bundles, is not mixed in)."""
import ctypes as C

import numpy as np
import pytest

from ephemeris_amd import _lib
from ephemeris_amd.layout import GP_DTYPE, MASS_DTYPE, RING_DTYPE
from ephemeris_amd.system import GpuSystems
from tests.util import GOLDEN, bits_equal, load_npz

pytestmark = pytest.mark.gpu


H2D, D2H = 1, 2       # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    _lib.lib()                                    # loads the library and with it its HIP runtime
    h = C.CDLL("libamdhip64.so.7")                # the runtime already loaded (matched by soname)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class _Dev:
    """a device copy of a numpy array (or n zero bytes)"""

    def __init__(self, h, a=None, nbytes=0):
        self.h = h
        self.nbytes = a.nbytes if a is not None else nbytes
        self.ptr = C.c_void_p()
        assert h.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        if a is not None:
            assert h.hipMemcpy(self.ptr, np.ascontiguousarray(a).ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0
        else:
            assert h.hipMemset(self.ptr, 0, self.nbytes) == 0

    def read(self, dtype):
        host = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.h.hipMemcpy(host.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H) == 0
        return host

    def free(self):
        assert self.h.hipFree(self.ptr) == 0


def _noalloc(d, summation):
    h = _hip()
    m = np.ascontiguousarray(d["masses_in"], MASS_DTYPE)
    gps = np.ascontiguousarray(d["gps"], GP_DTYPE)
    rings = np.ascontiguousarray(d["rings"], RING_DTYPE)
    dm = _Dev(h, m)
    dg = _Dev(h, gps) if len(gps) else None
    dr = _Dev(h, rings) if len(rings) else None
    out = _Dev(h, nbytes=m.nbytes)
    rc = _lib.lib().eph_accel_noalloc(dm.ptr, len(m), dg.ptr if dg else None, len(gps), dr.ptr if dr else None,
                                      len(rings), summation, out.ptr)
    _lib.check(rc)
    res = out.read(MASS_DTYPE)
    for b in (dm, dg, dr, out):
        if b:
            b.free()
    return res


def _resident(d, summation):
    g = GpuSystems(d["masses_in"], None, d["gps"], d["rings"])
    g.accel(summation)
    m, _ = g.download()
    g.close()
    return m


@pytest.mark.parametrize("first,second", [(("full504", "pool"), ("small25", "serial")),
                                          (("solar504", "pool"), ("full504", "serial"))])
def test_noalloc_equals_resident_path_twice(first, second):
    _lib.init()
    for name, tag in (first, second):
        print(f"\n[case] eph_accel_noalloc {name} {tag}", flush=True)
        d = load_npz(GOLDEN / f"accel_{name}_{tag}.npz")
        summ = _lib.SUMMATION_POOL if tag == "pool" else _lib.SUMMATION_SERIAL
        a = _noalloc(d, summ)
        b = _resident(d, summ)
        bad = [f for f in MASS_DTYPE.names if not bits_equal(a[f], b[f])]
        print(f"{len(a)} bodies: fields differing from eph_accel: {bad}", flush=True)
        assert not bad


def test_noalloc_rejects_sizes_beyond_its_limits():
    _lib.init()
    h = _hip()
    d = load_npz(GOLDEN / "accel_small25_serial.npz")
    m = np.ascontiguousarray(d["masses_in"], MASS_DTYPE)
    dm = _Dev(h, m)
    L = _lib.lib()
    assert L.eph_accel_noalloc(dm.ptr, 1025, None, 0, None, 0, 0, dm.ptr) == _lib.EPH_EINVAL
    assert b"1024" in L.eph_last_error()
    host = m.copy()
    assert L.eph_accel_noalloc(host.ctypes.data_as(C.c_void_p), len(m), None, 0, None, 0, 0,
                               host.ctypes.data_as(C.c_void_p)) == _lib.EPH_EINVAL
    assert b"device memory" in L.eph_last_error()
    dm.free()
