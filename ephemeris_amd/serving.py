"""Mixed serving workload: Poisson arrivals of test-particle jobs and
ensemble jobs, driven through the server's worker thread.

Job mix (defaults): 70% particle jobs with log-uniform counts in
[1e2, 1e5] and spans uniform in 1..365 days at dt = 1 day; 30% ensemble jobs
of 1..32 clones of the 504-body system truncated to 64..504 bodies with
spans of 1..30 days at dt = 300 s (whole days: the ensemble quantum is one
day, 288 steps).
"""
from __future__ import annotations

import tempfile
import time
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional

import numpy as np

from . import _lib
from . import generator as gen
from .ensemble import load_members, write_members


@dataclass
class JobSpec:
    kind: str                 # "particles" | "ensemble"
    arrival: float            # seconds after the start
    days: int
    n: int = 0                # particles
    members: Optional[List[int]] = None   # indices into the clone pool


def make_mix(n_jobs: int, rate: float, seed: int = 20254, pool_size: int = 64, frac_particles: float = 0.7,
             part_n=(1e2, 1e5), part_days=(1, 365), clones=(1, 32), ens_days=(1, 30)) -> List[JobSpec]:
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for _ in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        if rng.random() < frac_particles:
            n = int(round(np.exp(rng.uniform(np.log(part_n[0]), np.log(part_n[1])))))
            jobs.append(JobSpec("particles", t, int(rng.integers(part_days[0], part_days[1] + 1)), n=n))
        else:
            c = int(rng.integers(clones[0], clones[1] + 1))
            jobs.append(JobSpec("ensemble", t, int(rng.integers(ens_days[0], ens_days[1] + 1)),
                                members=[int(x) for x in rng.integers(0, pool_size, size=c)]))
    return jobs


class ClonePool:
    """Clones of the 504-body system (64..504 bodies, perturbed), each loaded
    through the product loader; jobs draw members from the pool."""

    def __init__(self, size=64, seed=20252, workdir=None):
        specs = gen.make_ensemble(size, seed=seed)
        workdir = Path(workdir or tempfile.mkdtemp(prefix="eph_pool_"))
        self.members = []
        for cfg in write_members(specs, workdir):
            masses, sys_n, gps, rings, teph = load_members([cfg])
            self.members.append((masses, gps, rings, teph))

    def job_arrays(self, idx):
        masses, gps, rings, teph, sys_n = [], [], [], [], []
        ngp = nring = 0
        for k in idx:
            m, g, r, t = self.members[k]
            m = m.copy()
            m["gp"] = np.where(m["gp"] >= 0, m["gp"] + ngp, -1)
            m["ring"] = np.where(m["ring"] >= 0, m["ring"] + nring, -1)
            ngp += len(g)
            nring += len(r)
            masses.append(m)
            gps.append(g)
            rings.append(r)
            teph.append(t)
            sys_n.append(len(m))
        return (np.concatenate(masses), np.array(sys_n, np.int64), np.concatenate(gps), np.concatenate(rings),
                np.concatenate(teph))


def serve(server, jobs: List[JobSpec], pool: ClonePool, particles, time_scale: float = 1.0, timeout=1200.0,
          progress=None):
    """Submit jobs at their arrival times (wall clock, scaled) while the
    server's worker runs; wait for all; return per-job records."""
    r_all, v_all = particles
    server.start()
    t0 = time.perf_counter()
    ids, offs = [], []
    off = 0
    for js in jobs:
        wait = js.arrival * time_scale - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        if js.kind == "particles":
            if off + js.n > len(r_all):
                off = 0
            jid = server.submit_particles(r_all[off:off + js.n], v_all[off:off + js.n], js.days)
            offs.append(off)
            off += js.n
        else:
            m, sn, g, r, t = pool.job_arrays(js.members)
            jid = server.submit_ensemble(m, sn, g, r, t, n_steps=js.days * 288)
            offs.append(-1)
        ids.append(jid)
        if progress:
            progress(f"submitted job {jid} ({js.kind}) at {time.perf_counter() - t0:.1f}s")
    deadline = time.perf_counter() + timeout
    while True:                              # wait, reporting progress every 10 s
        try:
            server.wait(min(10.0, max(deadline - time.perf_counter(), 0.01)))
            break
        except _lib.EphError as e:
            if e.code != _lib.EPH_EBUSY or time.perf_counter() > deadline:
                raise
            if progress:
                st = server.stats()
                progress(f"t={time.perf_counter() - t0:.0f}s running={st['jobs_running']} "
                         f"queued={st['jobs_queued']} done={st['jobs_done']} ticks={st['ticks']}")
    makespan = time.perf_counter() - t0
    server.stop()
    recs = []
    for js, jid, o in zip(jobs, ids, offs):
        i = server.info(jid)
        recs.append(dict(job=jid, kind=js.kind, days=js.days, units=i["units"], status=i["status_name"], offset=o,
                         latency=i["t_done"] - i["t_submit"], queued=i["t_admit"] - i["t_submit"],
                         error=i["error"]))
    return recs, makespan


def latency_summary(recs):
    lat = np.array([r["latency"] for r in recs if r["status"] == "done"])
    out = {"jobs": len(recs), "done": int(len(lat)),
           "failed": sum(r["status"] == "failed" for r in recs)}
    if len(lat):
        out.update(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                   mean_s=float(lat.mean()), max_s=float(lat.max()))
        if len(lat) < 100:
            # with fewer than 100 samples the 99th percentile is interpolated
            # between the two slowest jobs, so it is close to the maximum
            out["p99_note"] = f"{len(lat)} jobs: p99 lies between the two slowest latencies"
    for kind in ("particles", "ensemble"):
        k = np.array([r["latency"] for r in recs if r["status"] == "done" and r["kind"] == kind])
        if len(k):
            out[kind] = {"n": int(len(k)), "p50_s": float(np.percentile(k, 50)), "p99_s": float(np.percentile(k, 99))}
    return out
