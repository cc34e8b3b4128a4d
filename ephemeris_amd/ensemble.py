"""Ensembles of independent systems integrated as one GPU batch.

Each member is loaded through the product loader from the reference's text
formats (so an independent check can load the very same bytes), then all
members are concatenated into one resident eph_system: one launch sequence
per RK12 stage advances every member, and no tile, reduction or scratch row
ever mixes two members, so a member's result does not depend on what else is
in the batch.
"""
from __future__ import annotations

import tempfile
from pathlib import Path
from typing import List, Sequence

import numpy as np

from . import generator as gen
from .layout import GP_DTYPE, MASS_DTYPE, RING_DTYPE
from .system import GpuSystems, MSystem


def write_members(specs: Sequence[gen.SystemSpec], out_dir) -> List[Path]:
    out = Path(out_dir)
    return [gen.write_system(s, out / f"m{k:04d}", "Member") for k, s in enumerate(specs)]


def load_members(configs: Sequence[Path]):
    """Load each member config; return masses (concatenated, model indices
    rebased), sys_n, gps, rings and the members' epochs."""
    masses, sys_n, gps, rings, teph = [], [], [], [], []
    ngp = nring = 0
    for cfg in configs:
        m = MSystem(cfg, analyse=False)
        mm = m.masses
        g, r = m.models()
        mm["gp"] = np.where(mm["gp"] >= 0, mm["gp"] + ngp, -1)
        mm["ring"] = np.where(mm["ring"] >= 0, mm["ring"] + nring, -1)
        ngp += len(g)
        nring += len(r)
        masses.append(mm)
        gps.append(g)
        rings.append(r)
        sys_n.append(len(mm))
        teph.append(m.t_eph())
        m.close()
    return (np.concatenate(masses), np.array(sys_n, np.int64),
            np.concatenate(gps) if gps else np.zeros(0, GP_DTYPE),
            np.concatenate(rings) if rings else np.zeros(0, RING_DTYPE), np.array(teph))


class Ensemble:
    """A batch of member systems on the GPU."""

    def __init__(self, configs: Sequence[Path]):
        self.configs = list(configs)
        masses, self.sys_n, self.gps, self.rings, self.t0 = load_members(self.configs)
        self.offsets = np.concatenate([[0], np.cumsum(self.sys_n)])
        self.gpu = GpuSystems(masses, self.sys_n, self.gps, self.rings, self.t0)

    @classmethod
    def synthetic(cls, n_clones=256, seed=20252, workdir=None, base=None):
        specs = gen.make_ensemble(n_clones, seed=seed, base=base)
        workdir = Path(workdir or tempfile.mkdtemp(prefix="eph_ens_"))
        return cls(write_members(specs, workdir))

    @property
    def n_members(self):
        return len(self.sys_n)

    def pairs_per_eval(self):
        n = self.sys_n.astype(np.int64)
        return int((n * (n - 1)).sum())

    def integrate(self, dt, n_step):
        self.gpu.integrate(dt, n_step)

    def member(self, masses, k):
        return masses[self.offsets[k]:self.offsets[k + 1]]

    def download(self):
        return self.gpu.download()
