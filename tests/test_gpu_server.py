"""Serving loop: every job's result in the mix equals, bit for bit, the same
job run alone through the plain entry points; a job that goes non-finite
fails alone; admission respects the memory budget and slots."""
import tempfile

import numpy as np
import pytest

from ephemeris_amd import generator as gen
from ephemeris_amd.ensemble import load_members, write_members
from ephemeris_amd.particles import ParticleSet
from ephemeris_amd.server import DONE, FAILED, Server
from ephemeris_amd.system import GpuSystems, MSystem
from tests.util import bits_equal

pytestmark = pytest.mark.gpu

ENS_DT, ENS_Q = 300.0, 6
PART_DT, PART_Q, PART_MAX = 86400.0, 3, 24


@pytest.fixture(scope="module")
def setup():
    f9 = gen.make_field9(gen.make_full())
    cfg = gen.write_system(f9, tempfile.mkdtemp(), "Field9")
    specs = gen.make_ensemble(4, seed=7, n_lo=64, n_hi=130)
    members = [load_members([c]) for c in write_members(specs, tempfile.mkdtemp())]
    r, v = gen.make_particles(700, seed=11, field9=f9)
    return cfg, members, r, v


def _field(cfg):
    m = MSystem(cfg, analyse=False)
    gps, rings = m.models()
    f = GpuSystems(m.masses, None, gps, rings, [m.t_eph()])
    m.close()
    return f


def _job(members, idx):
    ms, sn, gs, rs, ts = [], [], [], [], []
    ngp = nring = 0
    for k in idx:
        m, _, g, r, t = members[k]
        m = m.copy()
        m["gp"] = np.where(m["gp"] >= 0, m["gp"] + ngp, -1)
        m["ring"] = np.where(m["ring"] >= 0, m["ring"] + nring, -1)
        ngp += len(g)
        nring += len(r)
        ms.append(m), gs.append(g), rs.append(r), ts.append(t), sn.append(len(m))
    return np.concatenate(ms), np.array(sn), np.concatenate(gs), np.concatenate(rs), np.concatenate(ts)


def _alone_ensemble(job, n_steps):
    m, sn, g, r, t = job
    s = GpuSystems(m, sn, g, r, t)
    for _ in range(n_steps // ENS_Q):
        s.integrate(ENS_DT, ENS_Q)
    return s.download()


def test_results_alone_equal_results_in_mix(setup):
    cfg, members, r, v = setup
    print("\n[case] server mix: 3 ensemble jobs (one poisoned), 3 particle jobs", flush=True)
    srv = Server(_field(cfg), ens_dt=ENS_DT, ens_quantum=ENS_Q, part_dt=PART_DT, part_quantum=PART_Q,
                 part_max_steps=PART_MAX)
    jobs_e = {}
    jobs_p = {}
    e1 = _job(members, [0])
    jobs_e[srv.submit_ensemble(*e1, n_steps=12)] = (e1, 12)
    jobs_p[srv.submit_particles(r[:300], v[:300], 7)] = (slice(0, 300), 7)
    srv.tick()
    e2 = _job(members, [1, 2])
    jobs_e[srv.submit_ensemble(*e2, n_steps=18)] = (e2, 18)
    jobs_p[srv.submit_particles(r[300:350], v[300:350], 20)] = (slice(300, 350), 20)
    bad = _job(members, [3])
    bad[0]["naccel_x"][5] = np.nan         # admitted (state finite), turns non-finite in the first stage
    bad_id = srv.submit_ensemble(*bad, n_steps=12)
    srv.tick()
    srv.tick()
    jobs_p[srv.submit_particles(r[350:700], v[350:700], 5)] = (slice(350, 700), 5)
    srv.run_until_idle()
    st = srv.stats()
    print("stats:", st, flush=True)

    info = srv.info(bad_id)
    assert info["status"] == FAILED and "non-finite" in info["error"], info
    for jid, (job, n) in jobs_e.items():
        i = srv.info(jid)
        assert i["status"] == DONE, i
        m, t = srv.result_ensemble(jid)
        ma, ta = _alone_ensemble(job, n)
        bad_f = [f for f in m.dtype.names if not bits_equal(m[f], ma[f])]
        print(f"ensemble job {jid}: {len(m)} bodies, {n} steps, mismatching fields {bad_f}", flush=True)
        assert not bad_f
        assert bits_equal(t, ta)
    for jid, (sl, n) in jobs_p.items():
        i = srv.info(jid)
        assert i["status"] == DONE, i
        pr, pv, pe = srv.result_particles(jid)
        ps = ParticleSet(_field(cfg), r[sl], v[sl])
        ps.integrate(PART_DT, n)
        ar, av, ae = ps.download()
        print(f"particle job {jid}: {len(pr)} particles, {n} steps", flush=True)
        assert bits_equal(pr, ar) and bits_equal(pv, av) and bits_equal(pe, ae)
    assert st["jobs_done"] == 5 and st["jobs_failed"] == 1
    srv.close()


def test_admission_by_memory_budget_and_worker(setup):
    cfg, members, r, v = setup
    print("\n[case] server budget: jobs queue behind a small memory budget, worker thread", flush=True)
    one = _job(members, [0])
    need = 535 * 8 * (len(one[0]) + 64) + 16384
    srv = Server(_field(cfg), ens_dt=ENS_DT, ens_quantum=ENS_Q, part_dt=PART_DT, part_quantum=PART_Q,
                 part_max_steps=PART_MAX, mem_budget=int(need * 1.5))
    ids = [srv.submit_ensemble(*one, n_steps=6) for _ in range(3)]
    srv.tick()
    infos = [srv.info(j) for j in ids]
    assert [i["status"] for i in infos].count(DONE) == 1      # one at a time fits the budget
    srv.start()
    srv.wait(300)
    srv.stop()
    infos = [srv.info(j) for j in ids]
    assert all(i["status"] == DONE for i in infos)
    admits = sorted(i["t_admit"] for i in infos)
    assert admits[0] < admits[1] < admits[2]
    res = [srv.result_ensemble(j)[0] for j in ids]
    assert bits_equal(res[0], res[1]) and bits_equal(res[1], res[2])
    with pytest.raises(Exception):
        srv.submit_ensemble(*_job(members, [0, 1, 2]), n_steps=6)   # larger than the whole budget
    with pytest.raises(Exception):
        srv.submit_ensemble(*one, n_steps=7)                        # not a multiple of the quantum
    srv.close()


def test_mixed_kinds_wait_under_a_tight_budget(setup):
    """A job that fits the budget alone waits while a job of the other kind
    holds part of it, and runs once that job is done."""
    cfg, members, r, v = setup
    print("\n[case] server budget: ensemble job queued behind a particle job", flush=True)
    one = _job(members, [0])
    probe = Server(_field(cfg), ens_dt=ENS_DT, ens_quantum=ENS_Q, part_dt=PART_DT, part_quantum=PART_Q,
                   part_max_steps=PART_MAX)
    pb = probe.info(probe.submit_particles(r[:600], v[:600], PART_MAX))["bytes"]
    eb = probe.info(probe.submit_ensemble(*one, n_steps=6))["bytes"]
    probe.close()
    budget = max(pb, eb) + min(pb, eb) // 2                  # each alone fits, both together do not
    srv = Server(_field(cfg), ens_dt=ENS_DT, ens_quantum=ENS_Q, part_dt=PART_DT, part_quantum=PART_Q,
                 part_max_steps=PART_MAX, mem_budget=int(budget))
    pj = srv.submit_particles(r[:600], v[:600], PART_MAX)
    srv.tick()
    assert srv.info(pj)["status"] not in (FAILED,)
    ej = srv.submit_ensemble(*one, n_steps=6)
    srv.tick()
    st_e = srv.info(ej)["status"]
    print(f"budget {budget} B (particle job {pb} B, ensemble job {eb} B); ensemble status after one tick: {st_e}",
          flush=True)
    assert st_e != FAILED
    srv.run_until_idle()
    assert srv.info(pj)["status"] == DONE and srv.info(ej)["status"] == DONE
    m, t = srv.result_ensemble(ej)
    ma, ta = _alone_ensemble(one, 6)
    assert bits_equal(m, ma) and bits_equal(t, ta)
    srv.close()


def test_snapshot_checkpoint_resume(setup, tmp_path):
    """A running job's snapshot is its state after the quanta done so far; a
    member written from it as a checkpoint, loaded again and integrated for
    the rest of the span ends where the served job ends."""
    cfg, members, r, v = setup
    print("\n[case] server snapshot: queued, running after one quantum, done; checkpoint resume", flush=True)
    srv = Server(_field(cfg), ens_dt=ENS_DT, ens_quantum=ENS_Q, part_dt=PART_DT, part_quantum=PART_Q,
                 part_max_steps=PART_MAX)
    other = _job(members, [0])
    srv.submit_ensemble(*other, n_steps=2 * ENS_Q)          # ahead of the job in the resident batch
    job = _job(members, [1, 2])
    jid = srv.submit_ensemble(*job, n_steps=3 * ENS_Q)
    m0, t0, k0 = srv.snapshot_ensemble(jid)
    assert k0 == 0 and bits_equal(m0, job[0]) and bits_equal(t0, job[4])
    srv.tick()
    m1, t1, k1 = srv.snapshot_ensemble(jid)
    ma, ta = _alone_ensemble(job, ENS_Q)
    assert k1 == ENS_Q
    assert not [f for f in m1.dtype.names if not bits_equal(m1[f], ma[f])]
    assert bits_equal(t1, ta)
    # the first member of the job as a checkpoint, resumed for the remaining steps
    n0 = int(job[1][0])
    ckpt = tmp_path / "member0.ckpt"
    w = MSystem.from_masses(m1[:n0], job[2], job[3], tuple(t1[0]))
    w.analyse(True)                                         # a checkpoint holds the orbit structure
    w.save_checkpoint(ckpt)
    w.close()
    back = MSystem.from_checkpoint(ckpt)
    gps, rings = back.models()
    resumed = GpuSystems(back.masses, None, gps, rings, [back.t_eph()])
    back.close()
    for _ in range(2):
        resumed.integrate(ENS_DT, ENS_Q)
    mr, tr = resumed.download()
    srv.run_until_idle()
    m3, t3, k3 = srv.snapshot_ensemble(jid)
    mf, tf = srv.result_ensemble(jid)
    assert k3 == 3 * ENS_Q and bits_equal(m3, mf) and bits_equal(t3, tf)
    pj = srv.submit_particles(r[:10], v[:10], PART_Q)
    for bad_id in (pj, 10 ** 9):                            # a particle job, an unknown id
        srv._shape.setdefault(bad_id, srv._shape[jid])
        with pytest.raises(Exception):
            srv.snapshot_ensemble(bad_id)
    keep = [f for f in mf.dtype.names if f not in ("gp", "ring")]
    bad_f = [f for f in keep if not bits_equal(mr[f], mf[:n0][f])]
    print(f"resumed member: {n0} bodies, mismatching fields {bad_f}", flush=True)
    assert not bad_f and bits_equal(tr[0], tf[0])
    srv.close()


def test_snapshots_while_the_worker_runs(setup):
    """Snapshots taken from another thread while the worker ticks: each
    holds whole quanta (steps done a multiple of the quantum, never
    decreasing) and equals the job run alone for that many steps."""
    import time
    cfg, members, r, v = setup
    print("\n[case] server snapshot from the caller's thread while the worker runs", flush=True)
    srv = Server(_field(cfg), ens_dt=ENS_DT, ens_quantum=ENS_Q, part_dt=PART_DT, part_quantum=PART_Q,
                 part_max_steps=PART_MAX)
    job = _job(members, [2])
    total = 40 * ENS_Q
    jid = srv.submit_ensemble(*job, n_steps=total)
    srv.start()
    snaps = {}
    t_end = time.monotonic() + 120
    while time.monotonic() < t_end:
        m, t, k = srv.snapshot_ensemble(jid)
        snaps.setdefault(k, (m, t))
        if k >= total or srv.info(jid)["status"] in (DONE, FAILED):
            break
        time.sleep(0.002)
    srv.wait(120)
    srv.stop()
    ks = list(snaps)
    print("steps seen in snapshots:", ks, flush=True)
    assert srv.info(jid)["status"] == DONE
    assert ks == sorted(ks) and all(k % ENS_Q == 0 for k in ks)
    mid = [k for k in ks if 0 < k < total]
    for k in mid[:2]:
        ma, ta = _alone_ensemble(job, k)
        assert bits_equal(snaps[k][0], ma) and bits_equal(snaps[k][1], ta), k
    mf, tf = srv.result_ensemble(jid)
    m3, t3, k3 = srv.snapshot_ensemble(jid)
    assert k3 == total and bits_equal(m3, mf) and bits_equal(t3, tf)
    srv.close()
