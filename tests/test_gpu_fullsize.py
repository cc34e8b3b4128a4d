"""Parity at the full benchmark sizes, through properties that do not need
the oracle to run at that size: every member of the 256-clone ensemble
(N ~ U(64, 504), the XCD-dealt four-lane kernels) and every particle of a
1e6-particle set is the same, bit for bit, as the same system or particle
integrated alone, which the smaller tests check against the oracle."""
import numpy as np
import pytest

from ephemeris_amd import generator as gen
from ephemeris_amd.ensemble import Ensemble, load_members
from ephemeris_amd.particles import ParticleSet
from ephemeris_amd.system import GpuSystems, MSystem
from tests.util import bits_equal

pytestmark = pytest.mark.gpu

DT = 300.0


def test_ensemble256_members_equal_alone():
    print("\n[case] 256-clone ensemble, 2 steps; members 0, 97, 255 alone", flush=True)
    ens = Ensemble.synthetic(256)
    assert ens.n_members == 256 and ens.sys_n.sum() > 60000
    ens.integrate(DT, 2)
    print("batch kernel variant", ens.gpu.kernel_variant(), flush=True)
    m_all, t_all = ens.download()
    for k in (0, 97, 255):
        masses, sys_n, gps, rings, t0 = load_members([ens.configs[k]])
        alone = GpuSystems(masses, sys_n, gps, rings, t0)
        alone.integrate(DT, 2)
        ma, ta = alone.download()
        mb = ens.member(m_all, k)
        bad = [f for f in ma.dtype.names if f not in ("gp", "ring") and not bits_equal(ma[f], mb[f])]
        print(f"member {k}: {len(ma)} bodies, mismatching fields {bad}", flush=True)
        assert not bad
        assert bits_equal(ta[0], t_all[k])


def test_particles_1e6_sample_equal_alone():
    print("\n[case] 1e6 particles, 1 daily step; 500 sampled particles alone", flush=True)
    f9 = gen.make_field9(gen.make_full())
    import tempfile
    cfg = gen.write_system(f9, tempfile.mkdtemp(prefix="eph_f9_"), "Field9")
    r, v = gen.make_particles(1000000, field9=f9)

    def field():
        m = MSystem(cfg, analyse=False)
        gps, rings = m.models()
        f = GpuSystems(m.masses, None, gps, rings, [m.t_eph()])
        m.close()
        return f

    big = ParticleSet(field(), r, v)
    big.integrate(86400.0, 1)
    br, bv, be = big.download()
    idx = np.sort(np.random.default_rng(5).choice(len(r), 500, replace=False))
    small = ParticleSet(field(), r[idx], v[idx])
    small.integrate(86400.0, 1)
    sr, sv, se = small.download()
    assert bits_equal(br[idx], sr) and bits_equal(bv[idx], sv) and bits_equal(be[idx], se)
    assert np.all(np.isfinite(br))
