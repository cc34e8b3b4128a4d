"""Massless test particles in the field of a small system (eph_particles).

The field system (N <= 48, e.g. a star and its planets) is integrated by
one persistent workgroup that records, stage by stage, the partner data its
bodies present to others; one thread per particle then replays those records.
Each particle evolves exactly as it would as one more body appended to the
field system in the reference (single-thread summation order), and the field
is unaffected by the particles."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import generator as gen
from .system import GpuSystems, MSystem, _ptr


class ParticleSet:
    def __init__(self, field: GpuSystems, r, v, GM=gen.PARTICLE_GM, R=gen.PARTICLE_R_KM * 1000.0):
        r = np.ascontiguousarray(r, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        if r.shape != v.shape:
            raise ValueError("r and v must have the same shape")
        self.field = field
        self.n = len(r)
        h = C.c_void_p()
        _lib.check(_lib.lib().eph_particles_create(field.h, self.n, _ptr(r), _ptr(v), float(GM), float(R), C.byref(h)))
        self.h = h

    @classmethod
    def from_config(cls, field_config, r, v, **kw):
        m = MSystem(field_config, analyse=False)
        gps, rings = m.models()
        field = GpuSystems(m.masses, None, gps, rings, [m.t_eph()])
        m.close()
        return cls(field, r, v, **kw)

    def integrate(self, dt, n_step):
        _lib.check(_lib.lib().eph_particles_integrate(self.h, float(dt), int(n_step)))

    def download(self):
        r = np.zeros((self.n, 6))
        v = np.zeros((self.n, 6))
        ex = np.zeros((self.n, 2))
        _lib.check(_lib.lib().eph_particles_download(self.h, _ptr(r), _ptr(v), _ptr(ex)))
        return r, v, ex

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().eph_particles_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
