"""Python handle on a batch of systems resident in HBM (eph_system)."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from .layout import GP_DTYPE, MASS_DTYPE, RING_DTYPE


def _ptr(a):
    return None if a is None or len(a) == 0 else a.ctypes.data_as(C.c_void_p)


class GpuSystems:
    """One or more independent systems integrated together on the GPU.

    ``masses`` holds the bodies of all systems back to back (``sys_n`` gives
    the count of each); model indices in ``masses['gp']``/``['ring']`` index
    ``gps``/``rings``.  ``t_eph`` is an (n_sys, 2) array of double-double
    epochs."""

    def __init__(self, masses: np.ndarray, sys_n: Sequence[int] = None, gps=None, rings=None,
                 t_eph=None, device=0):
        _lib.init(device)
        masses = np.ascontiguousarray(masses, dtype=MASS_DTYPE)
        sys_n = np.ascontiguousarray([len(masses)] if sys_n is None else sys_n, dtype=np.int64)
        gps = np.ascontiguousarray(np.zeros(0, GP_DTYPE) if gps is None else gps, dtype=GP_DTYPE)
        rings = np.ascontiguousarray(np.zeros(0, RING_DTYPE) if rings is None else rings, dtype=RING_DTYPE)
        if t_eph is None:
            t_eph = np.zeros((len(sys_n), 2))
        t_eph = np.ascontiguousarray(t_eph, dtype=np.float64).reshape(len(sys_n), 2)
        if int(sys_n.sum()) != len(masses):
            raise ValueError("sum(sys_n) must equal len(masses)")
        h = C.c_void_p()
        _lib.check(_lib.lib().eph_system_create(len(sys_n), _ptr(sys_n), _ptr(masses), _ptr(gps), len(gps),
                                                _ptr(rings), len(rings), _ptr(t_eph), C.byref(h)))
        self.h = h
        self.sys_n = sys_n
        self.n_sys = len(sys_n)
        self.nb = len(masses)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().eph_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download(self):
        m = np.zeros(self.nb, MASS_DTYPE)
        t = np.zeros((self.n_sys, 2))
        _lib.check(_lib.lib().eph_system_download(self.h, _ptr(m), _ptr(t)))
        return m, t

    def upload(self, masses, t_eph=None):
        masses = np.ascontiguousarray(masses, dtype=MASS_DTYPE)
        t = None if t_eph is None else np.ascontiguousarray(t_eph, dtype=np.float64)
        _lib.check(_lib.lib().eph_system_upload(self.h, _ptr(masses), _ptr(t)))

    def deform(self):
        _lib.check(_lib.lib().eph_deform(self.h))

    def accel(self, summation=_lib.SUMMATION_POOL):
        _lib.check(_lib.lib().eph_accel(self.h, summation))

    def integrate(self, dt, n_step, summation=_lib.SUMMATION_POOL):
        _lib.check(_lib.lib().eph_integrate_ex(self.h, float(dt), int(n_step), summation))

    def set_variant(self, variant=0, allow_persistent=True):
        """0 automatic, 1 one thread per body, 2 one wavefront per body (non-point
        pass fused where it fits), 3 one wavefront per body with the non-point
        pass in its own kernels, 4 four lanes per body."""
        _lib.check(_lib.lib().eph_system_set_variant(self.h, int(variant), int(allow_persistent)))

    def set_autotune(self, enable=True):
        """Time the kernel variants on a device copy at the first integrate (default on)."""
        _lib.check(_lib.lib().eph_system_set_autotune(self.h, int(bool(enable))))

    def kernel_variant(self):
        return _lib.check(_lib.lib().eph_system_kernel_variant(self.h))

    def set_waves(self, waves_per_body=0):
        """Wavefronts per body of variants 2 and 3: 0 automatic, 1, 2 or 4."""
        _lib.check(_lib.lib().eph_system_set_waves(self.h, int(waves_per_body)))

    def kernel_waves(self):
        return _lib.check(_lib.lib().eph_system_kernel_waves(self.h))

    def set_graph(self, enable=True):
        """Replay each RK12 step as one captured HIP graph (default on)."""
        _lib.check(_lib.lib().eph_system_set_graph(self.h, int(bool(enable))))

    def profile(self, enable=True):
        _lib.check(_lib.lib().eph_profile_enable(self.h, int(enable)))

    def profile_classes(self, names):
        """Time only the named kernel classes (the names profile_read returns)."""
        known = list(self.profile_read())
        mask = 0
        for nm in names:
            mask |= 1 << known.index(nm)
        _lib.check(_lib.lib().eph_profile_classes(self.h, mask))

    def profile_reset(self):
        _lib.check(_lib.lib().eph_profile_reset(self.h))

    def profile_read(self):
        names = (C.c_char_p * 16)()
        avg = (C.c_double * 16)()
        cnt = (C.c_int64 * 16)()
        n = _lib.check(_lib.lib().eph_profile_read(self.h, names, avg, cnt, 16))
        return {names[i].decode(): (avg[i], cnt[i]) for i in range(n)}


class MSystem:
    """A system in the msystem interface of the reference: loaded from the
    reference's text inputs or a checkpoint, or built from bodies in memory;
    accel / integrate on the CPU back end or on the GPU (parallel_option 1 /
    USE_GPU 1), combined integration, orbit-structure analysis."""

    CPU_RK12, GPU_RK12, COMBINED_RK12 = 0, 1, 2

    def __init__(self, config_path=None, summation_at_load=_lib.SUMMATION_SERIAL, *, handle=None, threads=0,
                 analyse=True):
        h = C.c_void_p() if handle is None else handle
        if handle is None:
            _lib.check(_lib.lib().eph_msystem_load_ex(str(config_path).encode(), summation_at_load, int(bool(analyse)),
                                                      C.byref(h)))
        self.h = h
        if threads:
            self.set_threads(threads)

    @classmethod
    def from_masses(cls, masses, gps=None, rings=None, t_eph=(0.0, 0.0), threads=0):
        masses = np.ascontiguousarray(masses, dtype=MASS_DTYPE)
        gps = np.ascontiguousarray(np.zeros(0, GP_DTYPE) if gps is None else gps, dtype=GP_DTYPE)
        rings = np.ascontiguousarray(np.zeros(0, RING_DTYPE) if rings is None else rings, dtype=RING_DTYPE)
        t = (C.c_double * 2)(*t_eph)
        h = C.c_void_p()
        _lib.check(_lib.lib().eph_msystem_create(len(masses), _ptr(masses), _ptr(gps), len(gps), _ptr(rings),
                                                 len(rings), t, C.byref(h)))
        return cls(handle=h, threads=threads)

    @classmethod
    def from_checkpoint(cls, path, threads=0):
        h = C.c_void_p()
        _lib.check(_lib.lib().eph_msystem_load_checkpoint(str(path).encode(), C.byref(h)))
        return cls(handle=h, threads=threads)

    def save_checkpoint(self, path):
        _lib.check(_lib.lib().eph_msystem_save_checkpoint(self.h, str(path).encode()))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().eph_msystem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return int(_lib.lib().eph_msystem_size(self.h))

    @property
    def masses(self):
        m = np.zeros(len(self), MASS_DTYPE)
        _lib.check(_lib.lib().eph_msystem_masses(self.h, _ptr(m), len(m)))
        return m

    def set_masses(self, masses):
        masses = np.ascontiguousarray(masses, dtype=MASS_DTYPE)
        _lib.check(_lib.lib().eph_msystem_set_masses(self.h, _ptr(masses), len(masses)))

    def models(self):
        ng, nr = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().eph_msystem_models(self.h, None, 0, None, 0, C.byref(ng), C.byref(nr)))
        gps, rings = np.zeros(ng.value, GP_DTYPE), np.zeros(nr.value, RING_DTYPE)
        _lib.check(_lib.lib().eph_msystem_models(self.h, _ptr(gps), len(gps), _ptr(rings), len(rings), None, None))
        return gps, rings

    def t_eph(self):
        t = (C.c_double * 2)()
        _lib.check(_lib.lib().eph_msystem_time(self.h, t))
        return t[0], t[1]

    def set_t_eph(self, t):
        _lib.check(_lib.lib().eph_msystem_set_time(self.h, (C.c_double * 2)(*t)))

    def params(self):
        p = (C.c_double * 8)()
        integ = C.c_int64()
        _lib.check(_lib.lib().eph_msystem_params(self.h, p, C.byref(integ)))
        names = ("delta_t", "data_cadence", "max_ephm_length", "combined_delta_t", "GM_max_child",
                 "GM_max_parent", "GM_max_tiny", "Period_max_child")
        d = dict(zip(names, p))
        d["integrator"] = integ.value
        return d

    def accel(self, parallel_option=0):
        """1: GPU; 0: CPU threads (thread-pool order when N is large); -1: CPU single thread."""
        if parallel_option > 0:
            _lib.init()
        _lib.check(_lib.lib().eph_msystem_accel(self.h, int(parallel_option)))

    def integrate(self, dt, n_step, use_gpu=1):
        """USE_GPU 1: GPU RK12; 0: CPU RK12."""
        if use_gpu:
            _lib.init()
        _lib.check(_lib.lib().eph_msystem_integrate(self.h, float(dt), int(n_step), int(use_gpu)))

    def combined_integrate(self, dt, n_combine, n_step, use_gpu=1):
        if use_gpu:
            _lib.init()
        _lib.check(_lib.lib().eph_msystem_combined_integrate(self.h, float(dt), int(n_combine), int(n_step),
                                                             int(use_gpu)))

    def update(self, t):
        _lib.check(_lib.lib().eph_msystem_update(self.h, float(t)))

    def analyse(self, reconstruct=False):
        t = (C.c_double * 2)()
        _lib.check(_lib.lib().eph_msystem_analyse(self.h, int(bool(reconstruct)), t))
        return t[0], t[1]

    def analysis_ok(self):
        """False when the last orbit-structure analysis did not settle."""
        t = (C.c_double * 2)()
        return _lib.lib().eph_msystem_analyse(self.h, 0, t) == 0

    def barycens(self):
        """[(pid, hid, gid, tid, mid, children, GM_sys)] of the orbit hierarchy."""
        n = _lib.check(_lib.lib().eph_msystem_barycens(self.h, None, None, None, 0, 0))
        ids = np.zeros((max(n, 1), 7), np.int64)
        gm = np.zeros((max(n, 1), 2))
        nch = n * n + 1
        ch = np.zeros(nch, np.int64)
        _lib.check(_lib.lib().eph_msystem_barycens(self.h, _ptr(ids), _ptr(gm), _ptr(ch), n, nch))
        return [(int(a[0]), int(a[1]), int(a[2]), int(a[3]), int(a[4]),
                 [int(c) for c in ch[a[6]:a[6] + a[5]]], gm[i, 0] + gm[i, 1]) for i, a in enumerate(ids[:n])]

    def scale(self, mid, factor):
        _lib.check(_lib.lib().eph_msystem_scale(self.h, int(mid), float(factor)))

    def scale_geopotential(self, mid, factor):
        _lib.check(_lib.lib().eph_msystem_scale_geopotential(self.h, int(mid), float(factor)))

    def scale_ring(self, mid, factor):
        _lib.check(_lib.lib().eph_msystem_scale_ring(self.h, int(mid), float(factor)))

    def set_tidal(self, parent, matrix, children):
        m = np.ascontiguousarray(matrix, np.float64).reshape(9)
        ch = np.ascontiguousarray(children, np.int64)
        _lib.check(_lib.lib().eph_msystem_set_tidal(self.h, int(parent), _ptr(m), _ptr(ch), len(ch)))

    def set_threads(self, n):
        _lib.check(_lib.lib().eph_msystem_set_threads(self.h, int(n)))

    def get_mid(self, ssid):
        return int(_lib.lib().eph_msystem_get_mid(self.h, ssid.encode()))

    def diagnostics(self):
        n = _lib.check(_lib.lib().eph_msystem_diagnostics(self.h, None, None, None, None, 0))
        body = (C.c_int64 * max(n, 1))()
        kind = (C.c_int * max(n, 1))()
        val = (C.c_double * max(n, 1))()
        lim = (C.c_double * max(n, 1))()
        _lib.check(_lib.lib().eph_msystem_diagnostics(self.h, body, kind, val, lim, n))
        return [(body[i], ("collision", "timestep")[kind[i]], val[i], lim[i]) for i in range(n)]
