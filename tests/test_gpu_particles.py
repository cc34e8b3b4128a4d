"""Test particles: each particle's trajectory must equal (bit for bit) the
oracle's trajectory of the same particle carried as one more body of the
field system, and the field must be unaffected by the particles."""
import tempfile
from dataclasses import replace

import numpy as np
import pytest

from ephemeris_amd import generator as gen
from ephemeris_amd.particles import ParticleSet
from ephemeris_amd.system import GpuSystems, MSystem
from oracle import pyoracle as po
from tests.util import bits_equal

pytestmark = pytest.mark.gpu

DT = 86400.0


@pytest.fixture(scope="module")
def field9():
    base = gen.make_full()
    f9 = gen.make_field9(base)
    d = tempfile.mkdtemp()
    return f9, gen.write_system(f9, d, "Field9")


def _oracle_with_particle(f9, r, v):
    spec = gen.SystemSpec(list(f9.bodies) + [gen.particle_body(r, v)], delta_t=DT)
    d = tempfile.mkdtemp()
    return po.OSystem.load(gen.write_system(spec, d, "FieldP"), pool_at_load=False)


def test_particles_match_oracle_bitwise(field9):
    f9, cfg = field9
    r, v = gen.make_particles(300, seed=5, field9=f9)
    print("\n[case] 300 particles, 40 daily steps", flush=True)
    ps = ParticleSet.from_config(cfg, r, v)
    ps.integrate(DT, 40)
    pr, pv, ex = ps.download()
    fm, _ = ps.field.download()
    for i in (0, 17, 299):
        o = _oracle_with_particle(f9, r[i], v[i])
        o.integrate(DT, 40, pool=True)
        m = o.masses
        ref_r = np.array([[m[f"r_{c}_hi"][9], m[f"r_{c}_lo"][9]] for c in "xyz"]).ravel()
        ref_v = np.array([[m[f"v_{c}_hi"][9], m[f"v_{c}_lo"][9]] for c in "xyz"]).ravel()
        dr = np.abs((pr[i, 0::2] + pr[i, 1::2]) - (ref_r[0::2] + ref_r[1::2])).max()
        print(f"particle {i}: max |dr| = {dr:.3e} m", flush=True)
        assert bits_equal(pr[i], ref_r) and bits_equal(pv[i], ref_v), i
        # the field in the oracle's 10-body run equals the GPU field integrated alone
        for f in ("r_x_hi", "r_x_lo", "v_y_hi", "v_y_lo", "r_z_lo"):
            assert bits_equal(fm[f], m[f][:9]), f


def test_particles_independent_of_batch_composition(field9):
    f9, cfg = field9
    r, v = gen.make_particles(1000, seed=6, field9=f9)
    a = ParticleSet.from_config(cfg, r, v)
    a.integrate(DT, 10)
    ra, _, _ = a.download()
    b = ParticleSet.from_config(cfg, r[500:507], v[500:507])
    b.integrate(DT, 10)
    rb, _, _ = b.download()
    assert bits_equal(ra[500:507], rb)


def test_field_small_path_matches_generic(field9):
    """The persistent one-workgroup path and the per-stage kernels agree."""
    f9, cfg = field9
    m = MSystem(cfg, analyse=False)
    m.integrate(DT, 30)
    a = m.masses
    o = po.OSystem.load(cfg, pool_at_load=False)
    o.integrate(DT, 30, pool=True)
    assert bits_equal(a["r_x_hi"], o.masses["r_x_hi"]) and bits_equal(a["r_y_lo"], o.masses["r_y_lo"])
