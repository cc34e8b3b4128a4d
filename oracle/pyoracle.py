"""ORACLE -- test infrastructure only.

ctypes binding of the C restatement in eph_oracle.c.  Imported by tests/,
by the smoke entry point and by the CPU-baseline leg of bench.py, and by
nothing else.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"


def _v(name):
    return [(name + "_" + c, "f8") for c in "xyz"]


def _ddv(name):
    return [(f"{name}_{c}_{p}", "f8") for c in "xyz" for p in ("hi", "lo")]


def _m(name):
    return [(f"{name}_{a}{c}", "f8") for a in "xyz" for c in "xyz"]


def _ddm(name):
    return [(f"{name}_{a}{c}_{p}", "f8") for a in "xyz" for c in "xyz" for p in ("hi", "lo")]


# field layout of o_mass (eph_oracle.h)
MASS_DTYPE = np.dtype(
    _ddv("r") + _ddv("v") + _ddv("GL") + _ddm("s") + _v("w")
    + [("GM", "f8"), ("exJ2", "f8"), ("sid", "u8")]
    + [(k, "f8") for k in ("GM0", "dGM", "dJ2", "R", "inertia", "k2", "k2r", "tide_delay",
                           "tide_delay_factor", "lum", "recpt", "A", "R2", "sR2", "rR2G_4c")]
    + _m("C_static") + [("gp", "i8"), ("ring", "i8"), ("phi", "f8"), ("beta2", "f8")]
    + _v("naccel") + _v("daccel") + _v("dtorque") + _v("gaccel") + _v("beta")
    + _m("GI") + _m("C_potential")
    + _v("Egrad") + _v("idaccel") + _v("idtorque") + _v("jdaccel") + _v("jdtorque")
    + [("min_distance", "f8"), ("max_influence", "f8")])

GP_DTYPE = np.dtype([("Nz", "i8"), ("Nt", "i8"), ("c", "f8", (77,))])
RING_DTYPE = np.dtype([("N", "i8"), ("GM_ratio", "f8"), ("A", "f8"), ("J2", "f8"),
                       ("GL_R2", "f8"), ("GsR2", "f8", (64,)), ("R_R", "f8", (64,)),
                       ("H_R", "f8", (64,))])


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P, D, I64, I = C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_int
        L.o_load_config.restype = P
        L.o_load_config.argtypes = [C.c_char_p, I]
        L.o_load_error.restype = C.c_char_p
        L.o_clone.restype = P
        L.o_clone.argtypes = [P]
        L.o_subset.restype = P
        L.o_subset.argtypes = [P, C.POINTER(C.c_int64), I64]
        L.o_free.argtypes = [P]
        L.o_size.restype = I64
        L.o_size.argtypes = [P]
        L.o_masses.restype = P
        L.o_masses.argtypes = [P]
        L.o_set_threads.argtypes = [P, I]
        L.o_accel.argtypes = [P, I, I]
        L.o_deform.argtypes = [P]
        L.o_integrate.argtypes = [P, C.c_double, I64, I]
        L.o_update.argtypes = [P, C.c_double]
        L.o_scale.argtypes = [P, I64, C.c_double]
        L.o_scale_geopotential.argtypes = [P, I64, C.c_double]
        L.o_scale_ring.argtypes = [P, I64, C.c_double]
        L.o_dd_ops.argtypes = [D, D, D]
        L.o_dd_sqrt.argtypes = [D, D]
        L.o_rotation_matrix.argtypes = [D, C.c_double, D]
        L.o_inverse.argtypes = [D, D]
        L.o_normalize.restype = C.c_double
        L.o_normalize.argtypes = [D]
        L.o_asc_node.argtypes = [D, D]
        L.o_gp_sum.argtypes = [P, C.c_double, D, D]
        L.o_ring_sum.argtypes = [P, C.c_double, D, D]
        L.o_fzcorr.restype = C.c_float
        L.o_fzcorr.argtypes = [C.c_float, C.c_float, I]
        L.o_deform_by.argtypes = [P, I64, I64]
        L.o_rk12_table.restype = D
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OSystem:
    """A loaded system in the oracle (msystem of the reference)."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("oracle load failed: " + lib().o_load_error().decode())
        self.h = handle

    @classmethod
    def load(cls, config_path, pool_at_load=False):
        return cls(lib().o_load_config(str(config_path).encode(), int(pool_at_load)))

    def clone(self):
        return OSystem(lib().o_clone(self.h))

    def subset(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        return OSystem(lib().o_subset(self.h, idx.ctypes.data_as(C.POINTER(C.c_int64)), len(idx)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_free(self.h)
            self.h = None

    def __len__(self):
        return int(lib().o_size(self.h))

    @property
    def masses(self) -> np.ndarray:
        """A numpy view (no copy) of the o_mass array."""
        n = len(self)
        buf = (C.c_char * (MASS_DTYPE.itemsize * n)).from_address(lib().o_masses(self.h))
        return np.frombuffer(buf, dtype=MASS_DTYPE, count=n)

    def set_threads(self, n):
        return lib().o_set_threads(self.h, int(n))

    def accel(self, parallel_option=0, pool=True):
        lib().o_accel(self.h, int(parallel_option), int(pool))

    def deform(self):
        lib().o_deform(self.h)

    def integrate(self, dt, n_step, pool=True):
        lib().o_integrate(self.h, float(dt), int(n_step), int(pool))

    def scale(self, mid, factor):
        lib().o_scale(self.h, int(mid), float(factor))

    def scale_geopotential(self, mid, factor):
        lib().o_scale_geopotential(self.h, int(mid), float(factor))

    def scale_ring(self, mid, factor):
        lib().o_scale_ring(self.h, int(mid), float(factor))

    def mid(self, ssid):
        key = int.from_bytes(ssid.encode().ljust(8, b"\0"), "little")
        hits = np.nonzero(self.masses["sid"] == key)[0]
        return int(hits[0]) if len(hits) else -1

    def raw_models(self):
        """(gp pointer, n_gp, ring pointer, n_ring) of the o_system."""
        p = C.cast(self.h, C.POINTER(C.c_int64))
        return p[7], p[5], p[8], p[6]

    def models(self):
        """Copies of the geopotential and ring tables (GP_DTYPE, RING_DTYPE)."""
        gp, ngp, rg, nrg = self.raw_models()
        gps = (np.frombuffer((C.c_char * (GP_DTYPE.itemsize * ngp)).from_address(gp), GP_DTYPE).copy()
               if ngp else np.zeros(0, GP_DTYPE))
        rings = (np.frombuffer((C.c_char * (RING_DTYPE.itemsize * nrg)).from_address(rg), RING_DTYPE).copy()
                 if nrg else np.zeros(0, RING_DTYPE))
        return gps, rings

    def t_eph(self):
        p = C.cast(self.h, C.POINTER(C.c_double))
        # o_system: int64 n; o_mass* m; o_dd t_eph ...
        return p[2] + p[3]


def rk12_table() -> np.ndarray:
    p = lib().o_rk12_table()
    return np.ctypeslib.as_array(p, shape=(325, 2)).copy()
