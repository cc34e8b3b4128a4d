"""Write the golden input systems and run the oracle over them.

Inputs (committed as text in the reference formats):
  tests/golden/small25/   25-body Newtonian+1PN+J2 system
  tests/golden/full504/   504-body full-model system

Trajectory fixtures (committed as .npz of float64 arrays, no pickle):
  traj_small25.npz   state after 288 steps of 300 s (1 day)
  traj_full504_<steps>.npz   state after 288, 2880, 8640, 28800 and
                     105192 steps (1, 10, 30, 100 and 365.25 days)

Protocol (mirrors the production driver): load without a thread pool, so the
initial accel takes the single-thread branch (mass_loader.cpp:555); then call
integrate(300, chunk) with the pool present, one call per checkpoint
interval.  The GPU tests replay the same call sequence.

    python tests/golden/make_trajectories.py inputs
    python tests/golden/make_trajectories.py small
    python tests/golden/make_trajectories.py full [--threads 6] [--upto STEPS]
"""
import argparse
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

FULL_CHECKPOINTS = [288, 2880, 8640, 28800, 105192]
SMALL_STEPS = 288
DT = 300.0

STATE_FIELDS = (["r_%s_%s" % (c, p) for c in "xyz" for p in ("hi", "lo")]
                + ["v_%s_%s" % (c, p) for c in "xyz" for p in ("hi", "lo")]
                + ["GL_%s_%s" % (c, p) for c in "xyz" for p in ("hi", "lo")]
                + ["s_%s%s_%s" % (a, c, p) for a in "xyz" for c in "xyz" for p in ("hi", "lo")]
                + ["w_%s" % c for c in "xyz"]
                + ["GM", "exJ2", "min_distance", "max_influence"]
                + ["naccel_%s" % c for c in "xyz"] + ["gaccel_%s" % c for c in "xyz"]
                + ["daccel_%s" % c for c in "xyz"] + ["dtorque_%s" % c for c in "xyz"])


def config_path(name):
    return HERE / name / ("Small_Config.txt" if name == "small25" else "Full_Config.txt")


def write_inputs():
    from ephemeris_amd import generator as g
    g.write_system(g.make_small(), HERE / "small25", "Small")
    g.write_system(g.make_full(), HERE / "full504", "Full")


def save_state(path, sys_, steps):
    m = sys_.masses
    np.savez_compressed(path, steps=np.int64(steps), t_eph=np.float64(sys_.t_eph()),
                        **{f: np.ascontiguousarray(m[f]) for f in STATE_FIELDS})


def run_small():
    from oracle import pyoracle as po
    s = po.OSystem.load(config_path("small25"), pool_at_load=False)
    s.integrate(DT, SMALL_STEPS, pool=True)
    save_state(HERE / "traj_small25.npz", s, SMALL_STEPS)


def run_full(threads, upto):
    from oracle import pyoracle as po
    s = po.OSystem.load(config_path("full504"), pool_at_load=False)
    s.set_threads(threads)
    done = 0
    t0 = time.time()
    for ck in FULL_CHECKPOINTS:
        if ck > upto:
            break
        s.integrate(DT, ck - done, pool=True)
        done = ck
        save_state(HERE / ("traj_full504_%d.npz" % ck), s, ck)
        print("checkpoint %d steps, %.0f s elapsed" % (ck, time.time() - t0), flush=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["inputs", "small", "full"])
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--upto", type=int, default=FULL_CHECKPOINTS[-1])
    a = ap.parse_args()
    if a.what == "inputs":
        write_inputs()
    elif a.what == "small":
        run_small()
    else:
        run_full(a.threads, a.upto)
