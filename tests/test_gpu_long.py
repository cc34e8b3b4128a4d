"""Long-horizon parity of the 504-body system against the oracle.

The oracle trajectory (tests/golden/make_trajectories.py) loads the system,
then calls integrate(300 s, chunk) once per checkpoint interval; the GPU run
replays the same call sequence, because the epoch used for the stage times
is taken at the start of each call.  Bar: max relative position difference
<= 1e-12 at every checkpoint (1 simulated year is 105,192 steps).

One GPU system is carried through the checkpoints in order, one test per
checkpoint, so the year ends in the last test (about two minutes of GPU
time after the 100-day checkpoint).  A test run on its own first replays the
earlier intervals.  Beside the position and velocity differences each
checkpoint reports the drift of linear and angular momentum (with spin) of
the GPU trajectory against the oracle's, both relative to the epoch value.
"""
import numpy as np
import pytest

from ephemeris_amd.system import GpuSystems
from tests.util import GOLDEN, load_npz, momenta, position, velocity

pytestmark = pytest.mark.gpu

CHECKPOINTS = [288, 2880, 8640, 28800, 105192]
POS_TOL = 1e-12


class _Carry:
    """The GPU system and the number of steps it has taken."""
    g = None
    done = 0
    p0 = None


def _advance_to(ck):
    if _Carry.g is None or _Carry.done > ck:
        d = load_npz(GOLDEN / "accel_full504_pool.npz")
        _Carry.g = GpuSystems(d["masses_in"], None, d["gps"], d["rings"])
        _Carry.done = 0
        _Carry.p0 = momenta(d["masses_in"])
    for c in CHECKPOINTS:
        if _Carry.done < c <= ck:
            _Carry.g.integrate(300.0, c - _Carry.done)
            _Carry.done = c
    return _Carry.g


@pytest.mark.timeout(900)
@pytest.mark.parametrize("ck", CHECKPOINTS)
def test_full504_checkpoint(ck):
    print(f"\n[case] full504 to {ck} steps ({ck / 288:.2f} days)", flush=True)
    g = _advance_to(ck)
    ref = load_npz(GOLDEN / f"traj_full504_{ck}.npz")
    m, t = g.download()
    r, rr = position(m), position(ref)
    dr = np.max(np.linalg.norm(r - rr, axis=1) / np.linalg.norm(rr, axis=1))
    dv = np.max(np.linalg.norm(velocity(m) - velocity(ref), axis=1) /
                np.maximum(np.linalg.norm(velocity(ref), axis=1), 1e-30))
    da = np.max(np.abs(r - rr))
    p0 = _Carry.p0
    pg, po = momenta(m), momenta(ref)
    drift = {k: (np.linalg.norm(pg[k] - p0[k]) / np.linalg.norm(p0[k]),
                 np.linalg.norm(po[k] - p0[k]) / np.linalg.norm(p0[k])) for k in p0}
    print(f"full504 after {ck} steps: max rel dr={dr:.3e} dv={dv:.3e} max |dr|={da:.3e} m; "
          + "; ".join(f"{k} drift GPU {a:.3e} oracle {b:.3e}" for k, (a, b) in drift.items())
          + f"; t_eph equal={t[0].sum() == float(ref['t_eph'])}", flush=True)
    assert t[0].sum() == float(ref["t_eph"])
    assert dr <= POS_TOL, dr
