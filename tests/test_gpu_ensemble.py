"""Batched ensembles: every member of a batch matches the same member
integrated alone (bit for bit) and the oracle on the same text inputs."""
import numpy as np
import pytest

from ephemeris_amd import ensemble, generator as gen
from ephemeris_amd.system import GpuSystems
from oracle import pyoracle as po
from tests.util import bits_equal, position

pytestmark = pytest.mark.gpu

# model indices are rebased inside a batch; every other field must match
STATE_FIELDS = [f for f in __import__("ephemeris_amd").MASS_DTYPE.names if f not in ("gp", "ring")]


@pytest.fixture(scope="module")
def members(tmp_path_factory):
    base = gen.make_full()
    specs = gen.make_ensemble(6, seed=99, base=base, n_lo=40, n_hi=200)
    return ensemble.write_members(specs, tmp_path_factory.mktemp("ens"))


def test_member_alone_equals_member_in_batch(members):
    print(f"\n[case] ensemble of {len(members)} members, 4 steps", flush=True)
    ens = ensemble.Ensemble(members)
    ens.integrate(300.0, 4)
    mb, tb = ens.download()
    for k in (0, 3, 5):
        alone = ensemble.Ensemble([members[k]])
        alone.integrate(300.0, 4)
        ma, ta = alone.download()
        bad = [f for f in STATE_FIELDS if not bits_equal(ens.member(mb, k)[f], ma[f])]
        assert not bad, (k, bad)
        assert tb[k].tolist() == ta[0].tolist()


def test_members_match_oracle(members):
    print("\n[case] ensemble members vs oracle, 4 steps", flush=True)
    ens = ensemble.Ensemble(members)
    ens.integrate(300.0, 4)
    mb, _ = ens.download()
    for k in (1, 4):
        o = po.OSystem.load(members[k], pool_at_load=False)
        o.integrate(300.0, 4, pool=True)
        mine, ref = ens.member(mb, k), o.masses
        dr = np.max(np.linalg.norm(position(mine) - position(ref), axis=1) / np.linalg.norm(position(ref), axis=1))
        print(f"member {k} (N={len(ref)}): max rel dr = {dr:.3e}", flush=True)
        assert dr <= 1e-13


def test_autotuned_batch_matches_fixed_variant(tmp_path_factory):
    """The autotuner (on a batch above its size floor) only picks a kernel
    variant; the trajectory bits equal those of a fixed variant."""
    print("\n[case] autotuned 16-member batch vs variant 1, 3 steps", flush=True)
    specs = gen.make_ensemble(16, seed=98, n_lo=150, n_hi=300)
    cfgs = ensemble.write_members(specs, tmp_path_factory.mktemp("tune"))
    masses, sys_n, gps, rings, t0 = ensemble.load_members(cfgs)
    assert sys_n.sum() >= 2048
    tuned = GpuSystems(masses, sys_n, gps, rings, t0)
    fixed = GpuSystems(masses, sys_n, gps, rings, t0)
    fixed.set_variant(1)
    tuned.integrate(300.0, 3)
    fixed.integrate(300.0, 3)
    print("autotuner chose variant", tuned.kernel_variant(), "wavefronts per body", tuned.kernel_waves(), flush=True)
    assert tuned.kernel_variant() in (1, 2, 4)
    assert tuned.kernel_waves() in ((1, 2, 4) if tuned.kernel_variant() == 2 else (0,))
    a, ta = tuned.download()
    b, tb = fixed.download()
    bad = [f for f in a.dtype.names if not bits_equal(a[f], b[f])]
    assert not bad and bits_equal(ta, tb), bad


def test_xcd_dealt_batch_matches_per_thread_variant(tmp_path_factory):
    """A batch of 24 systems on the four-lanes-per-body kernels, whose work
    lists are dealt to the XCDs by system and padded with empty items, gives
    the bits of the one-thread-per-body kernels."""
    print("\n[case] 24-member batch, variant 4 (XCD-dealt lists) vs variant 1, 2 steps", flush=True)
    specs = gen.make_ensemble(24, seed=97, n_lo=64, n_hi=200)
    cfgs = ensemble.write_members(specs, tmp_path_factory.mktemp("xcd"))
    masses, sys_n, gps, rings, t0 = ensemble.load_members(cfgs)
    k4 = GpuSystems(masses, sys_n, gps, rings, t0)
    pt = GpuSystems(masses, sys_n, gps, rings, t0)
    k4.set_variant(4)
    pt.set_variant(1)
    k4.integrate(300.0, 2)
    pt.integrate(300.0, 2)
    a, ta = k4.download()
    b, tb = pt.download()
    bad = [f for f in a.dtype.names if not bits_equal(a[f], b[f])]
    assert not bad and bits_equal(ta, tb), bad


def test_four_lane_batch_without_time_lag_or_items(tmp_path_factory):
    """Edge cases of the four-lane launches: a batch with no constant-time-lag
    body (the deformation launch has only four-lane blocks and item fields)
    and a batch with no non-point model (no item rows: no item fields in the
    deformation launch, no item sums in the main-force launch).  Both give
    the bits of the one-thread-per-body kernels."""
    specs = gen.make_ensemble(24, seed=95, n_lo=64, n_hi=200)
    cfgs = ensemble.write_members(specs, tmp_path_factory.mktemp("edge4"))
    masses, sys_n, gps, rings, t0 = ensemble.load_members(cfgs)
    assert (masses["tide_delay"] != 0).any() and len(gps) > 0 and len(rings) > 0
    no_lag = masses.copy()
    no_lag["tide_delay"] = 0.0
    no_models = masses.copy()
    no_models["gp"] = -1
    no_models["ring"] = -1
    cases = (("no constant-time-lag body", no_lag, gps, rings),
             ("no non-point model", no_models, gps[:0], rings[:0]))
    for label, m, g, r in cases:
        print(f"\n[case] 24-member batch with {label}: variant 4 vs variant 1, 2 steps", flush=True)
        k4 = GpuSystems(m, sys_n, g, r, t0)
        pt = GpuSystems(m, sys_n, g, r, t0)
        k4.set_variant(4)
        pt.set_variant(1)
        k4.integrate(300.0, 2)
        pt.integrate(300.0, 2)
        a, ta = k4.download()
        b, tb = pt.download()
        bad = [f for f in a.dtype.names if not bits_equal(a[f], b[f])]
        assert not bad and bits_equal(ta, tb), (label, bad)


def test_two_wavefront_fused_batch_matches_per_thread_variant(tmp_path_factory):
    """A batch of six large systems on the one-workgroup-per-body kernels:
    with more than 768 bodies the main force runs two wavefronts per body
    with the non-point pass fused (the own items' products formed one item
    at a time after the pair loop); bits equal the one-thread-per-body
    kernels'."""
    print("\n[case] 6-member batch of 300-504 bodies, variant 2 (two wavefronts, fused non-point) vs variant 1, 2 steps",
          flush=True)
    specs = gen.make_ensemble(6, seed=96, n_lo=300, n_hi=504)
    cfgs = ensemble.write_members(specs, tmp_path_factory.mktemp("wpb2"))
    masses, sys_n, gps, rings, t0 = ensemble.load_members(cfgs)
    assert sys_n.sum() > 768 and len(gps) > 0 and len(rings) > 0
    w2 = GpuSystems(masses, sys_n, gps, rings, t0)
    pt = GpuSystems(masses, sys_n, gps, rings, t0)
    w2.set_variant(2)
    pt.set_variant(1)
    w2.integrate(300.0, 2)
    pt.integrate(300.0, 2)
    a, ta = w2.download()
    b, tb = pt.download()
    bad = [f for f in a.dtype.names if not bits_equal(a[f], b[f])]
    assert not bad and bits_equal(ta, tb), bad
