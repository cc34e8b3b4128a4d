"""Benchmark of the force-evaluation + RK12 hot path on one MI355X.

Workload (N=1): the 504-body full-model system (seeded, tests/golden/full504),
RK12 at dt=300 s.  A step is one RK12 step = 25 force evaluations of the
system.  The headline is timed on the library's default path: K steps inside
one eph_integrate call, each step one replay of the captured HIP graph, with
the state already resident in HBM.  The roofline kernel is timed in a second
region of K steps launched from the host (graphs off) with HIP events around
the main-force launches on the library's stream.

Prints one JSON line:
  value   FP64 pair-interactions/s = K * 25 * N*(N-1) / wall time (graph replay)
  roofline  the main-force kernel: algorithmic FP64 FLOP per launch /
            average launch time measured with HIP events on the library's
            stream over the second timed region, against MI355X's FP64 vector
            peak (the kernels run on the FP64 vector ALUs, no MFMA)
  cpu_baseline  the oracle (C restatement of the reference CPU path, threaded
            over bodies with OpenMP) on this host's cores, on a bounded sample
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

FP64_VECTOR_PEAK_TFLOPS = 78.6
DT = 300.0
# algorithmic FP64 operations (add, sub, mul, div, sqrt, fma = 1 each) per
# ordered pair, counted from the kernel source (eph_kernels.hip)
FLOP_MAIN_PAIR = 294          # RELATIVITY 109, tidal (both quadrupoles) 92, Lense-Thirring 83, radiation 10
FLOP_DEFORM_PAIR = {"nk2": 54, "ndelay": 89, "damped": 217}
# double-double operations of eph_math.hpp: add/sub 11, mul 9, diff (dd
# difference rounded to double, per component) 12
DD_ADD, DD_MUL, DIFF3 = 11, 9, 36


def rk12_nonzero_rows():
    """Non-zero columns of each row k = 1..25 of the RK12 tableau."""
    import re
    rows = {}
    for line in (ROOT / "ephemeris_amd" / "tables" / "rk12_feagin.inc").read_text().splitlines():
        m = re.match(r"/\* k=\s*(\d+) \*/ (.*)", line)
        if m:
            ents = re.findall(r"\{([^,]+),([^}]+)\}", m.group(2))
            rows[int(m.group(1))] = [j for j, (a, b) in enumerate(ents)
                                     if float.fromhex(a.strip()) != 0 or float.fromhex(b.strip()) != 0]
    return rows


def particle_flop_per_step(nf, n_quad, n_spin):
    """Algorithmic FP64 operations of one RK12 step of one particle, counted
    from k_particles / particle_accel: per stage, the field bodies' potential
    and Newtonian pull (54 each), the reactions of the quadrupoles (88 each),
    RELATIVITY (108 each) and the partners' Lense-Thirring terms (31 each),
    plus 25 fixed; and the stage combination, which forms each stage
    velocity it uses from the stage sum kept when that stage was formed."""
    per_stage = nf * 54 + 16 + n_quad * 88 + nf * 108 + n_spin * 31 + 9
    rows = rk12_nonzero_rows()
    comb = 0
    for k in range(1, 26):
        for j in rows[k]:
            comb += 1 + 6                                          # fc, ca += fc * a_j
            comb += 3 * DD_MUL + 3 * DD_ADD                        # cv += c_kj * v_j (dd)
            if j > 0:
                comb += 3 + 3 * DD_ADD                             # v_j = v0 + dt * (kept sum_l a_jl a_l)
        comb += (3 * DD_MUL + 3 * DD_ADD) + (3 + 3 * DD_ADD)      # r_k, v_k
    return 25 * per_stage + comb


PROFILE = ROOT / "profiles" / "kernels_full504_ensemble256_particles1e6.md"


def profiled_traffic(kernel_prefix, grid):
    """Calibrated HBM bytes per dispatch (read + write) of a kernel at one grid
    size, from the committed rocprofv3 PMC summary (None if not listed)."""
    try:
        for line in PROFILE.read_text().splitlines():
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) == 6 and cols[0].strip("`").startswith(kernel_prefix) and cols[1] == str(grid):
                return (float(cols[4]) + float(cols[5])) * 1e6
    except (OSError, ValueError):
        pass
    return None


def load_workload():
    with np.load(ROOT / "tests" / "golden" / "accel_full504_pool.npz", allow_pickle=False) as z:
        return z["masses_in"], z["gps"], z["rings"]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(masses_n, sample_steps):
    from oracle import pyoracle as po                       # CPU baseline leg only
    s = po.OSystem.load(ROOT / "tests" / "golden" / "full504" / "Full_Config.txt", pool_at_load=False)
    cores = len(os.sched_getaffinity(0))            # the cores this process may run on
    s.set_threads(cores)
    t0 = time.perf_counter()
    s.integrate(DT, sample_steps, pool=True)
    el = time.perf_counter() - t0
    pairs = sample_steps * 25 * masses_n * (masses_n - 1)
    return {"value": pairs / el, "unit": "pair-interactions/s", "cores": cores, "cpu_model": cpu_model(),
            "threads": "OpenMP, one per core of the process's CPU affinity", "kind": "port",
            "sample": f"{sample_steps} RK12 steps of the 504-body system ({el:.1f} s)",
            "rk12_steps_per_s": sample_steps / el}


def batched_leg(steps, warmup, n_clones):
    """256 clones of the 504-body system (N ~ U(64,504)) in one batch."""
    from ephemeris_amd.ensemble import Ensemble
    t0 = time.perf_counter()
    ens = Ensemble.synthetic(n_clones)
    t_load = time.perf_counter() - t0
    ens.integrate(DT, warmup)
    g = ens.gpu
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    ens.integrate(DT, steps)
    el = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    variant = g.kernel_variant()
    pairs = ens.pairs_per_eval()
    main_ms = prof["main_force_pool"][0]
    achieved = pairs * FLOP_MAIN_PAIR / (main_ms * 1e-3) / 1e12
    # deformation pairs by tidal class (mass.impl:1-46): k2 = 0, k2 with no
    # time lag, constant time lag (one spin pass counted; those bodies take
    # 2-4, so the deform figure is a lower bound)
    m, _ = ens.download()
    cls_pairs = {"nk2": 0, "ndelay": 0, "damped": 0}
    for k in range(ens.n_members):
        mk = ens.member(m, k)
        nm = len(mk)
        nk2 = int((mk["k2"] == 0).sum())
        nd = int(((mk["k2"] != 0) & (mk["tide_delay"] != 0)).sum())
        cls_pairs["nk2"] += nk2 * (nm - 1)
        cls_pairs["damped"] += nd * (nm - 1)
        cls_pairs["ndelay"] += (nm - nk2 - nd) * (nm - 1)
    deform_ms = prof["deform_spin"][0]
    deform_flop = sum(cls_pairs[c] * FLOP_DEFORM_PAIR[c] for c in cls_pairs)
    deform_tf = deform_flop / (deform_ms * 1e-3) / 1e12
    return {"members": ens.n_members, "bodies": int(ens.sys_n.sum()), "pairs_per_eval": pairs,
            "steps": steps, "pair_interactions_per_s": steps * 25 * pairs / el,
            "rk12_member_steps_per_s": steps * ens.n_members / el, "ms_per_step": el / steps * 1e3,
            "load_s": t_load, "kernel_variant": {2: "one wavefront per body", 4: "four lanes per body"}.get(variant, variant),
            "roofline_main_force": {"achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                                    "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "avg_launch_ms": main_ms,
                                    "note": ("main-force FLOP over the whole launch, which also forms the non-point "
                                             "items' partial sums in extra workgroups (about 2% of its time)"
                                             if variant == 4 else "main-force FLOP over the main-force launch")},
            "roofline_deform": {"pairs_by_class": cls_pairs, "flop_per_pair": FLOP_DEFORM_PAIR,
                                "flop_first_pass": deform_flop, "ms_per_evaluation": deform_ms,
                                "achieved_lower_bound": deform_tf, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac_lower_bound": deform_tf / FP64_VECTOR_PEAK_TFLOPS,
                                "note": "the deformation launch (constant-time-lag bodies, the other bodies and the "
                                        "item fields); constant-time-lag bodies counted for one spin pass"},
            "kernel_ms": {k: round(v[0], 5) for k, v in prof.items()}}


def particles_leg(n_particles, steps):
    """Massless particles in the field of the star and 8 planets, dt = 1 day."""
    import tempfile
    from ephemeris_amd import generator as gen
    from ephemeris_amd.particles import ParticleSet
    f9 = gen.make_field9(gen.make_full())
    cfg = gen.write_system(f9, tempfile.mkdtemp(prefix="eph_f9_"), "Field9")
    r, v = gen.make_particles(n_particles, field9=f9)
    ps = ParticleSet.from_config(cfg, r, v)
    ps.integrate(86400.0, 1)
    g = ps.field
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    ps.integrate(86400.0, steps)
    el = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    nf = len(f9.bodies)
    fm, _ = g.download()
    n_quad = int((np.stack([fm[f"C_potential_{a}{b}"] for a in "xyz" for b in "xyz"], -1) != 0).any(-1).sum())
    n_spin = int((np.stack([fm[f"GL_{c}_hi"] for c in "xyz"], -1) != 0).any(-1).sum())
    flop_step = particle_flop_per_step(nf, n_quad, n_spin)
    part_ms = prof["particle_steps"][0] * prof["particle_steps"][1] / steps if "particle_steps" in prof else None
    roof = None
    if part_ms:
        tf = n_particles * flop_step / (part_ms * 1e-3) / 1e12
        roof = {"kernel": "k_particles", "flop_per_particle_step": flop_step, "quadrupole_field_bodies": n_quad,
                "spinning_field_bodies": n_spin, "ms_per_step": part_ms, "achieved": tf,
                "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP64_VECTOR_PEAK_TFLOPS}
    return {"particles": n_particles, "field_bodies": nf, "steps": steps,
            "particle_steps_per_s": n_particles * steps / el,
            "pair_interactions_per_s": n_particles * steps * 25 * nf / el, "ms_per_step": el / steps * 1e3,
            "kernel_ms": {k: round(v[0], 4) for k, v in prof.items() if v[1]}, "roofline": roof}


def serving_leg(n_jobs, rate, seed, verify):
    """Poisson arrivals of particle and ensemble jobs into the server."""
    import tempfile
    from ephemeris_amd import generator as gen
    from ephemeris_amd.particles import ParticleSet
    from ephemeris_amd.server import Server
    from ephemeris_amd.serving import ClonePool, latency_summary, make_mix, serve
    from ephemeris_amd.system import GpuSystems, MSystem
    from tests.util import bits_equal
    f9 = gen.make_field9(gen.make_full())
    cfg = gen.write_system(f9, tempfile.mkdtemp(prefix="eph_f9_"), "Field9")

    def field():
        m = MSystem(cfg, analyse=False)
        gps, rings = m.models()
        f = GpuSystems(m.masses, None, gps, rings, [m.t_eph()])
        m.close()
        return f

    t0 = time.perf_counter()
    pool = ClonePool(32)
    jobs = make_mix(n_jobs, rate, seed=seed, pool_size=32)
    npart = max([j.n for j in jobs if j.kind == "particles"] + [1])
    r, v = gen.make_particles(max(npart, 1000), seed=seed, field9=f9)
    srv = Server(field(), ens_quantum=288, part_quantum=16, part_max_steps=365)
    t_setup = time.perf_counter() - t0
    recs, makespan = serve(srv, jobs, pool, (r, v), progress=lambda m: print(m, file=sys.stderr, flush=True))
    st = srv.stats()
    out = {"jobs": n_jobs, "arrival_rate_per_s": rate, "seed": seed, "setup_s": t_setup, "makespan_s": makespan,
           "latency": latency_summary(recs),
           "pair_interactions_per_s": st["pair_interactions"] / makespan,
           "rk12_system_steps_per_s": st["rk12_system_steps"] / makespan,
           "particle_steps_per_s": st["particle_steps"] / makespan,
           "jobs_per_s": sum(x["status"] == "done" for x in recs) / makespan,
           "ticks": st["ticks"], "busy_s": st["busy_s"], "mem_peak_bytes": st["mem_peak"],
           "mem_budget_bytes": st["mem_budget"]}
    if verify:
        # the smallest job of each kind again, alone through the plain entry points
        checked = ok = 0
        pj = [(x, j) for x, j in zip(recs, jobs) if j.kind == "particles" and x["status"] == "done"]
        if pj:
            x, j = min(pj, key=lambda t: t[0]["units"])
            off = x["offset"]
            ps = ParticleSet(field(), r[off:off + j.n], v[off:off + j.n])
            ps.integrate(86400.0, j.days)
            a = ps.download()
            b = srv.result_particles(x["job"])
            checked += 1
            ok += all(bits_equal(p, q) for p, q in zip(a, b))
        ej = [(x, j) for x, j in zip(recs, jobs) if j.kind == "ensemble" and x["status"] == "done"]
        if ej:
            x, j = min(ej, key=lambda t: t[0]["units"] * t[0]["days"])
            m, sn, g, rg, t = pool.job_arrays(j.members)
            s = GpuSystems(m, sn, g, rg, t)
            for _ in range(j.days):
                s.integrate(300.0, 288)
            a, ta = s.download()
            b, tb = srv.result_ensemble(x["job"])
            checked += 1
            ok += bits_equal(a, b) and bits_equal(ta, tb)
        out["verified_alone_vs_mix"] = f"{ok}/{checked} bit-identical"
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=4)
    ap.add_argument("--clones", type=int, default=256)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--particles", type=int, default=1000000)
    ap.add_argument("--particle-steps", type=int, default=4)
    ap.add_argument("--no-particles", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=12)
    ap.add_argument("--serve-rate", type=float, default=1.0, help="job arrivals per second (Poisson)")
    ap.add_argument("--serve-seed", type=int, default=20254)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (calibrates HBM byte counters) and exit")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        x = torch.empty(a.copy_probe * 2 ** 20, dtype=torch.uint8, device="cuda:0")
        y = torch.empty_like(x)
        x.fill_(1)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel(), "copies": 5}))
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (each system, clone or particle "
              "set fits one MI355X; clones, jobs and particles would be split across GPUs by index with no "
              "exchange between them)", file=sys.stderr)
        sys.exit(2)

    from ephemeris_amd.system import GpuSystems
    masses, gps, rings = load_workload()
    n = len(masses)
    g = GpuSystems(masses, None, gps, rings)
    g.integrate(DT, max(a.warmup, 2))               # warm-up; captures the step graph
    pairs = a.steps * 25 * n * (n - 1)
    # timed region 1 (headline): the library's default path, one captured
    # graph replayed per RK12 step
    t0 = time.perf_counter()
    g.integrate(DT, a.steps)
    el = time.perf_counter() - t0
    value = pairs / el
    # timed region 2 (roofline): the same steps launched from the host, HIP
    # events around the launches of the roofline kernel (each timed launch
    # adds two event records)
    g.profile_classes(["main_force_pool"])
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    g.integrate(DT, a.steps)
    el_eager = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    # device time of every kernel class, outside the timed regions
    g.profile_classes(list(prof))
    g.profile(True)
    g.profile_reset()
    g.integrate(DT, 4)
    allprof = g.profile_read()
    g.profile(False)

    tot = {k: v[0] * v[1] for k, v in allprof.items()}
    dominant = max(tot, key=tot.get)
    main_ms = prof["main_force_pool"][0]
    flop_main = n * (n - 1) * FLOP_MAIN_PAIR
    achieved = flop_main / (main_ms * 1e-3) / 1e12
    roofline = {"kernel": "k_main_wpb<true, 4> (main force + fused non-point pass)", "kernel_class": "main_force_pool",
                "dominant_kernel": dominant, "bound": "fp64_valu",
                "compute": "FP64 vector ALU (78.6 TFLOP/s FP64 vector peak; no MFMA: not a dense contraction)",
                "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
                "traffic": profiled_traffic("void eph::k_main_wpb<true, 4>", n * 256), "traffic_unit": "bytes per launch",
                "flop_per_launch": flop_main, "avg_launch_ms": main_ms,
                "timed_region": "second region: host-launched steps with HIP events on the library stream",
                "kernel_ms": {k: round(v[0], 5) for k, v in allprof.items() if v[1]},
                "kernel_launches_timed_region": {k: v[1] for k, v in prof.items() if v[1]}}
    out = {"metric": "FP64 pair-interactions/s (504-body RK12)", "value": value,
           "unit": "pair-interactions/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": el / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (double-double state)", "data": "synthetic",
           "config": {"workload": "full504: 504-body full-model system, RK12 dt=300 s, plain single-system path",
                      "bodies": n, "parallelism": "single GPU", "launch": "one HIP graph replay per RK12 step"},
           "rk12_steps_per_s": a.steps / el, "roofline": roofline,
           "host_launched": {"rk12_steps_per_s": a.steps / el_eager, "ms_per_step": el_eager / a.steps * 1e3,
                             "pair_interactions_per_s": pairs / el_eager}}
    if not a.no_batch:
        b = batched_leg(a.batch_steps, 1, a.clones)
        b["vs_plain_pair_rate"] = b["pair_interactions_per_s"] / value
        out["batched"] = b
    if not a.no_particles:
        out["particles"] = particles_leg(a.particles, a.particle_steps)
    if not a.no_serve:
        out["serving"] = serving_leg(a.serve_jobs, a.serve_rate, a.serve_seed, verify=True)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
