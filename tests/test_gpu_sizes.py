"""Edge sizes against the oracle, through the C ABI.

* Systems of one, two and three bodies taken from the 504-body system: the
  star alone (no partner; it carries a geopotential, a luminosity and a mass
  rate), the star with the ringed planet (one partner each way, both
  non-point models), and the star, that planet and its first moon (the
  planet and the moon carry constant-time-lag tides).  The reference sums
  such systems in the single-thread order whatever the requested order
  (mass.cpp:216-226: one task), and so must the library.
* An 804-body system, the 504-body one with 300 more main-belt asteroids.
  It is above the 512-body limit of the fused non-point pass and of the
  deform kernel's cached pair geometry, so its non-point pass runs in its
  own kernels, its constant-time-lag bodies re-form the pair geometry on
  every spin pass, and its main force takes four rounds of 256 partners.
* 40 bodies with geopotentials, the ring and constant-time-lag tides, in the
  single-thread order on the one-workgroup-per-system path at 256 threads.
* A batch of the 1-, 2-, 3- and 804-body systems: each member ends with the
  bits it has alone.
* Empty inputs are rejected.

Force evaluations must equal the oracle's bit for bit (ring-carrying
outputs to 1e-13 relative: device libm), and a few RK12 steps must stay
within 1e-15 in position and 1e-13 in velocity."""
import numpy as np
import pytest

from ephemeris_amd import _lib
from ephemeris_amd import generator as gen
from ephemeris_amd.layout import ACCEL_OUTPUTS, MASS_DTYPE
from ephemeris_amd.system import GpuSystems
from oracle import pyoracle as po
from tests.util import CONFIGS, bits_equal, position, velocity

pytestmark = pytest.mark.gpu

RING_FIELDS = tuple(f"{q}_{c}" for q in ("daccel", "dtorque", "jdaccel", "jdtorque") for c in "xyz")


def _mismatch(m, ref):
    has_ring = bool((ref["ring"] >= 0).any())
    bad = {}
    for f in ACCEL_OUTPUTS:
        if bits_equal(m[f], ref[f]):
            continue
        a, b = np.asarray(m[f], np.float64), np.asarray(ref[f], np.float64)
        scale = np.maximum(np.abs(a), np.abs(b))
        rel = float(np.max(np.abs(a - b) / np.where(scale == 0, 1.0, scale)))
        if has_ring and f in RING_FIELDS and rel < 1e-13:
            continue
        bad[f] = rel
    return bad


def _traj(m, ref):
    r, rr = position(m), position(ref)
    v, vr = velocity(m), velocity(ref)
    dr = np.max(np.linalg.norm(r - rr, axis=1) / np.linalg.norm(rr, axis=1))
    dv = np.max(np.linalg.norm(v - vr, axis=1) / np.maximum(np.linalg.norm(vr, axis=1), 1e-30))
    return dr, dv


def _check_against_oracle(o, label, steps=3):
    """accel in the thread-pool order and `steps` RK12 steps, GPU vs oracle."""
    gps, rings = o.models()
    g = GpuSystems(o.masses.copy(), None, gps, rings)
    a = o.clone()
    a.accel(0, pool=True)
    g.accel(_lib.SUMMATION_POOL)
    m, _ = g.download()
    bad = _mismatch(m, a.masses.copy())
    print(f"{label}: accel fields beyond the allowance {bad}", flush=True)
    assert not bad
    g = GpuSystems(o.masses.copy(), None, gps, rings)
    b = o.clone()
    g.integrate(300.0, steps)
    b.integrate(300.0, steps, pool=True)
    m, _ = g.download()
    dr, dv = _traj(m, b.masses.copy())
    print(f"{label}: {steps} steps max rel dr={dr:.3e} dv={dv:.3e}", flush=True)
    assert dr <= 1e-15 and dv <= 1e-13


def test_systems_of_one_two_and_three_bodies():
    full = po.OSystem.load(CONFIGS["full504"], pool_at_load=False)
    ms = full.masses
    host = int(np.nonzero(ms["ring"] >= 0)[0][0])
    moon = host + 1
    assert ms["k2"][host] != 0 and ms["tide_delay"][moon] != 0 and ms["gp"][0] >= 0
    for idx in ([0], [0, host], [0, host, moon]):
        print(f"\n[case] {len(idx)}-body system {idx} of full504", flush=True)
        _check_against_oracle(full.subset(idx), f"{len(idx)} bodies", steps=4)


def test_forty_bodies_with_models_and_tides():
    """40 bodies of the 504-body system -- the star (geopotential), the
    planets (geopotentials, the ring) and the first moons (constant-time-lag
    tides): below 49 bodies the reference sums in the single-thread order, and
    the library integrates it on the one-workgroup-per-system path with 256
    threads.  Accel and four steps against the oracle."""
    print("\n[case] 40-body subset of full504 with models and tides, single-thread order", flush=True)
    full = po.OSystem.load(CONFIGS["full504"], pool_at_load=False)
    ms = full.masses
    idx = list(range(40))
    sub = ms[idx]
    assert (sub["gp"] >= 0).sum() >= 2 and (sub["ring"] >= 0).any() and (sub["tide_delay"] != 0).sum() >= 2
    _check_against_oracle(full.subset(idx), "40 bodies", steps=4)


def test_system_above_the_fused_and_cached_limits(tmp_path):
    print("\n[case] 804-body system (full504 + 300 asteroids): accel and 2 steps vs oracle", flush=True)
    spec = gen.add_asteroids(gen.make_full(), 300)
    cfg = gen.write_system(spec, tmp_path, "Large")
    o = po.OSystem.load(cfg, pool_at_load=False)
    assert len(o) == 804 and (o.masses["tide_delay"] != 0).sum() > 0 and (o.masses["ring"] >= 0).any()
    _check_against_oracle(o, "804 bodies", steps=2)
    # every kernel shape gives the same bits at this size too
    gps, rings = o.models()
    outs = []
    for variant, waves in ((1, 0), (2, 4), (3, 1), (4, 0)):
        g = GpuSystems(o.masses.copy(), None, gps, rings)
        g.set_variant(variant, allow_persistent=False)
        g.set_waves(waves)
        g.integrate(300.0, 2)
        outs.append(g.download()[0])
    for k in range(1, len(outs)):
        bad = [f for f in MASS_DTYPE.names if not bits_equal(outs[k][f], outs[0][f])]
        assert not bad, (k, bad)


def test_empty_inputs_are_rejected():
    m = po.OSystem.load(CONFIGS["small25"], pool_at_load=False).masses.copy()
    with pytest.raises(_lib.EphError):
        GpuSystems(m[:0], np.zeros(0, np.int64))
    with pytest.raises(_lib.EphError):
        GpuSystems(m, np.array([25, 0]))


def test_batch_of_extreme_sizes_equals_members_alone(tmp_path):
    """One batch holding systems of 1, 2, 3 and 804 bodies (single-thread and
    thread-pool orders side by side): every member's bits after three steps
    equal the member integrated alone."""
    print("\n[case] batch of 1-, 2-, 3- and 804-body systems vs each alone, 3 steps", flush=True)
    full = po.OSystem.load(CONFIGS["full504"], pool_at_load=False)
    ms = full.masses
    host = int(np.nonzero(ms["ring"] >= 0)[0][0])
    large = po.OSystem.load(gen.write_system(gen.add_asteroids(gen.make_full(), 300), tmp_path, "Large"),
                            pool_at_load=False)
    members = [full.subset([0]), full.subset([0, host]), full.subset([0, host, host + 1]), large]
    parts, sys_n, gps_all, rings_all = [], [], [], []
    ngp = nring = 0
    for o in members:
        m = o.masses.copy()
        g, r = o.models()
        m["gp"] = np.where(m["gp"] >= 0, m["gp"] + ngp, -1)
        m["ring"] = np.where(m["ring"] >= 0, m["ring"] + nring, -1)
        ngp += len(g)
        nring += len(r)
        parts.append(m), sys_n.append(len(m)), gps_all.append(g), rings_all.append(r)
    batch = GpuSystems(np.concatenate(parts), np.array(sys_n), np.concatenate(gps_all), np.concatenate(rings_all))
    batch.integrate(300.0, 3)
    mb, _ = batch.download()
    off = 0
    for o, n in zip(members, sys_n):
        g, r = o.models()
        alone = GpuSystems(o.masses.copy(), None, g, r)
        alone.integrate(300.0, 3)
        ma, _ = alone.download()
        mine = mb[off:off + n].copy()
        bad = [f for f in MASS_DTYPE.names if f not in ("gp", "ring") and not bits_equal(mine[f], ma[f])]
        print(f"{n} bodies: fields differing from the member alone {bad}", flush=True)
        assert not bad
        off += n
