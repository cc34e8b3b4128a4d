"""Serving loop over resident state (eph_server): continuous batching of
ensemble jobs (independent systems) and test-particle jobs (massless
particles in one shared field) on one GPU.

Jobs are admitted first-in first-out at tick boundaries while their HBM
footprint fits; every resident ensemble system advances in one batch and
every resident particle in one launch, and each job's result is the same
bits as the job run alone:

* ensemble job: ``GpuSystems(...).integrate(ens_dt, ens_quantum)`` repeated
  ``n_steps / ens_quantum`` times;
* particle job: ``ParticleSet(field copy, r, v).integrate(part_dt, n_steps)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import generator as gen
from .layout import GP_DTYPE, MASS_DTYPE, RING_DTYPE
from .system import GpuSystems, _ptr

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3
STATUS = {QUEUED: "queued", RUNNING: "running", DONE: "done", FAILED: "failed"}
ENSEMBLE, PARTICLES = 0, 1


class ServerConfig(C.Structure):
    _fields_ = [("mem_budget", C.c_int64), ("ens_dt", C.c_double), ("ens_quantum", C.c_int64),
                ("part_dt", C.c_double), ("part_quantum", C.c_int64), ("part_max_steps", C.c_int64),
                ("part_GM", C.c_double), ("part_R", C.c_double), ("max_systems", C.c_int64),
                ("max_particles", C.c_int64)]


class JobInfo(C.Structure):
    _fields_ = [("kind", C.c_int32), ("status", C.c_int32), ("steps_total", C.c_int64),
                ("steps_done", C.c_int64), ("bytes", C.c_int64), ("units", C.c_int64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("error", C.c_char * 160)]


class ServerStats(C.Structure):
    _fields_ = [("ticks", C.c_int64), ("jobs_submitted", C.c_int64), ("jobs_done", C.c_int64),
                ("jobs_failed", C.c_int64), ("jobs_queued", C.c_int64), ("jobs_running", C.c_int64),
                ("mem_used", C.c_int64), ("mem_peak", C.c_int64), ("mem_budget", C.c_int64),
                ("pair_interactions", C.c_double), ("rk12_system_steps", C.c_double),
                ("particle_steps", C.c_double), ("busy_s", C.c_double), ("t_now", C.c_double)]


def _struct_dict(s):
    return {name: (getattr(s, name).decode() if isinstance(getattr(s, name), bytes) else getattr(s, name))
            for name, _ in s._fields_}


class Server:
    def __init__(self, field: GpuSystems = None, *, mem_budget=0, ens_dt=300.0, ens_quantum=288,
                 part_dt=86400.0, part_quantum=16, part_max_steps=365, part_GM=gen.PARTICLE_GM,
                 part_R=gen.PARTICLE_R_KM * 1000.0, max_systems=0, max_particles=0):
        _lib.init()
        self.cfg = ServerConfig(int(mem_budget), float(ens_dt), int(ens_quantum), float(part_dt),
                                int(part_quantum), int(part_max_steps), float(part_GM), float(part_R),
                                int(max_systems), int(max_particles))
        h = C.c_void_p()
        _lib.check(_lib.lib().eph_server_create(C.byref(self.cfg), field.h if field is not None else None,
                                                C.byref(h)))
        self.h = h
        self._shape = {}

    # ---- submission ---------------------------------------------------------
    def submit_ensemble(self, masses, sys_n, gps=None, rings=None, t_eph=None, n_steps=None):
        masses = np.ascontiguousarray(masses, MASS_DTYPE)
        sys_n = np.ascontiguousarray(sys_n, np.int64)
        gps = np.ascontiguousarray(np.zeros(0, GP_DTYPE) if gps is None else gps, GP_DTYPE)
        rings = np.ascontiguousarray(np.zeros(0, RING_DTYPE) if rings is None else rings, RING_DTYPE)
        t = np.ascontiguousarray(np.zeros((len(sys_n), 2)) if t_eph is None else t_eph, np.float64)
        job = C.c_int64()
        _lib.check(_lib.lib().eph_server_submit_ensemble(
            self.h, len(sys_n), _ptr(sys_n), _ptr(masses), _ptr(gps), len(gps), _ptr(rings), len(rings),
            _ptr(t.reshape(-1)), int(n_steps), C.byref(job)))
        self._shape[job.value] = (ENSEMBLE, len(masses), len(sys_n))
        return job.value

    def submit_particles(self, r, v, n_steps):
        r = np.ascontiguousarray(r, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        if r.shape != v.shape:
            raise ValueError("r and v must have the same shape")
        job = C.c_int64()
        _lib.check(_lib.lib().eph_server_submit_particles(self.h, len(r), _ptr(r), _ptr(v), int(n_steps),
                                                          C.byref(job)))
        self._shape[job.value] = (PARTICLES, len(r), 0)
        return job.value

    # ---- driving --------------------------------------------------------------
    def tick(self):
        n = C.c_int64()
        _lib.check(_lib.lib().eph_server_tick(self.h, C.byref(n)))
        return n.value

    def run_until_idle(self, max_ticks=100000):
        for _ in range(max_ticks):
            self.tick()
            st = self.stats()
            if st["jobs_queued"] == 0 and st["jobs_running"] == 0:
                return
        raise RuntimeError("server did not drain")

    def start(self):
        _lib.check(_lib.lib().eph_server_start(self.h))

    def wait(self, timeout_s=600.0):
        _lib.check(_lib.lib().eph_server_wait(self.h, float(timeout_s)))

    def stop(self):
        _lib.check(_lib.lib().eph_server_stop(self.h))

    # ---- results --------------------------------------------------------------
    def info(self, job):
        i = JobInfo()
        _lib.check(_lib.lib().eph_server_job_info(self.h, int(job), C.byref(i)))
        d = _struct_dict(i)
        d["status_name"] = STATUS.get(d["status"], "?")
        return d

    def result_ensemble(self, job):
        kind, nb, ns = self._shape[job]
        m = np.zeros(nb, MASS_DTYPE)
        t = np.zeros((ns, 2))
        _lib.check(_lib.lib().eph_server_result_ensemble(self.h, int(job), _ptr(m), _ptr(t)))
        return m, t

    def snapshot_ensemble(self, job):
        """The job's systems at its last whole quantum and the steps done so far
        (its input while queued, its result once done)."""
        kind, nb, ns = self._shape[job]
        m = np.zeros(nb, MASS_DTYPE)
        t = np.zeros((ns, 2))
        steps = C.c_int64()
        _lib.check(_lib.lib().eph_server_snapshot_ensemble(self.h, int(job), _ptr(m), _ptr(t), C.byref(steps)))
        return m, t, steps.value

    def result_particles(self, job):
        kind, n, _ = self._shape[job]
        r = np.zeros((n, 6))
        v = np.zeros((n, 6))
        ex = np.zeros((n, 2))
        _lib.check(_lib.lib().eph_server_result_particles(self.h, int(job), _ptr(r), _ptr(v), _ptr(ex)))
        return r, v, ex

    def release(self, job):
        _lib.check(_lib.lib().eph_server_release(self.h, int(job)))
        self._shape.pop(job, None)

    def stats(self):
        s = ServerStats()
        _lib.check(_lib.lib().eph_server_stats_read(self.h, C.byref(s)))
        return _struct_dict(s)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().eph_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
