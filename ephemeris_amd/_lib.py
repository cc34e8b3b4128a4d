"""ctypes binding of lib/libephemeris_hip.so (C ABI: include/ephemeris_hip.h).

The HIP library is the only compute path: importing a function of this
module raises if the library has not been built."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libephemeris_hip.so"

EPH_OK, EPH_EINVAL, EPH_ENODEV, EPH_ENOMEM, EPH_EDEVICE, EPH_EIO, EPH_ESANITY, EPH_ENOTSUP, EPH_EBUSY = \
    0, -1, -2, -3, -4, -5, -6, -7, -8
SUMMATION_POOL, SUMMATION_SERIAL = 0, -1

# every symbol include/ephemeris_hip.h declares: (restype, argtypes)
P, I, I64, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
PP = C.POINTER(C.c_void_p)
SIGNATURES = {
    "eph_init": (I, [I]),
    "eph_last_error": (C.c_char_p, []),
    "eph_abi_version": (I, []),
    "eph_system_create": (I, [I64, P, P, P, I64, P, I64, P, PP]),
    "eph_system_destroy": (I, [P]),
    "eph_system_upload": (I, [P, P, P]),
    "eph_system_download": (I, [P, P, P]),
    "eph_system_bodies": (I64, [P]),
    "eph_system_count": (I64, [P]),
    "eph_deform": (I, [P]),
    "eph_accel": (I, [P, I]),
    "eph_integrate": (I, [P, D, I64]),
    "eph_integrate_ex": (I, [P, D, I64, I]),
    "eph_accel_noalloc": (I, [P, I64, P, I64, P, I64, I, P]),
    "eph_profile_enable": (I, [P, I]),
    "eph_profile_read": (I, [P, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.POINTER(C.c_int64), I]),
    "eph_profile_reset": (I, [P]),
    "eph_profile_classes": (I, [P, C.c_uint]),
    "eph_system_set_variant": (I, [P, I, I]),
    "eph_system_set_graph": (I, [P, I]),
    "eph_system_set_autotune": (I, [P, I]),
    "eph_system_kernel_variant": (I, [P]),
    "eph_system_set_waves": (I, [P, I]),
    "eph_system_kernel_waves": (I, [P]),
    "eph_particles_create": (I, [P, I64, P, P, D, D, PP]),
    "eph_particles_integrate": (I, [P, D, I64]),
    "eph_particles_download": (I, [P, P, P, P]),
    "eph_particles_count": (I64, [P]),
    "eph_particles_destroy": (I, [P]),
    "eph_server_create": (I, [P, P, PP]),
    "eph_server_submit_ensemble": (I, [P, I64, P, P, P, I64, P, I64, P, I64, C.POINTER(C.c_int64)]),
    "eph_server_submit_particles": (I, [P, I64, P, P, I64, C.POINTER(C.c_int64)]),
    "eph_server_tick": (I, [P, C.POINTER(C.c_int64)]),
    "eph_server_start": (I, [P]),
    "eph_server_wait": (I, [P, D]),
    "eph_server_stop": (I, [P]),
    "eph_server_job_info": (I, [P, I64, P]),
    "eph_server_result_ensemble": (I, [P, I64, P, P]),
    "eph_server_snapshot_ensemble": (I, [P, I64, P, P, C.POINTER(C.c_int64)]),
    "eph_server_result_particles": (I, [P, I64, P, P, P]),
    "eph_server_release": (I, [P, I64]),
    "eph_server_stats_read": (I, [P, P]),
    "eph_server_destroy": (I, [P]),
    "eph_msystem_load": (I, [C.c_char_p, I, PP]),
    "eph_msystem_load_ex": (I, [C.c_char_p, I, I, PP]),
    "eph_msystem_create": (I, [I64, P, P, I64, P, I64, P, PP]),
    "eph_msystem_load_checkpoint": (I, [C.c_char_p, PP]),
    "eph_msystem_save_checkpoint": (I, [P, C.c_char_p]),
    "eph_msystem_destroy": (I, [P]),
    "eph_msystem_size": (I64, [P]),
    "eph_msystem_masses": (I, [P, P, I64]),
    "eph_msystem_set_masses": (I, [P, P, I64]),
    "eph_msystem_models": (I, [P, P, I64, P, I64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "eph_msystem_time": (I, [P, C.POINTER(C.c_double)]),
    "eph_msystem_set_time": (I, [P, C.POINTER(C.c_double)]),
    "eph_msystem_params": (I, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "eph_msystem_accel": (I, [P, I]),
    "eph_msystem_integrate": (I, [P, D, I64, I]),
    "eph_msystem_combined_integrate": (I, [P, D, I64, I64, I]),
    "eph_msystem_update": (I, [P, D]),
    "eph_msystem_analyse": (I, [P, I, C.POINTER(C.c_double)]),
    "eph_msystem_barycens": (I64, [P, P, P, P, I64, I64]),
    "eph_msystem_scale": (I, [P, I64, D]),
    "eph_msystem_scale_geopotential": (I, [P, I64, D]),
    "eph_msystem_scale_ring": (I, [P, I64, D]),
    "eph_msystem_set_tidal": (I, [P, I64, P, P, I64]),
    "eph_msystem_set_threads": (I, [P, I]),
    "eph_msystem_get_mid": (I64, [P, C.c_char_p]),
    "eph_msystem_diagnostics": (I, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), I64]),
}


class EphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise EphError(rc, lib().eph_last_error().decode(errors="replace"))
    return rc


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(lib().eph_init(device))
        _initialised = True
