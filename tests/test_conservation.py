"""Conservation of linear and angular momentum, after the reference's own
test (src/tests/test_conservation.cpp:10-112): Saturn and four of its moons
from the reference's Solar System, Saturn's geopotential x3 and ring x1e7
(Saturn's own mass reduced so that the system's mass is unchanged), 6,000
RK12 steps of 900 s.  Momentum includes spin (GM GL) and, for the ring
host, the ring's share of the mass and its angular momentum.  Bars of the
reference: relative variation of either momentum < 1e-12, and the ring
correction factor < 1e-8.

The same run on the oracle, on the product's CPU back end (bit-identical to
the oracle) and on the GPU back end.  The last is marked gpu."""
import numpy as np
import pytest

from ephemeris_amd.system import MSystem
from oracle import pyoracle as po
from tests.util import CONFIGS

SIDS = ["699", "601", "603", "605", "606"]
STEPS = 6000
EPS_CONSERVATION = 1e-12          # TEST_EPSILON_CONSERVATION, test_conservation.cpp:6
EPS_RING = 1e-8                   # TEST_EPSILON_RING_CORRECTION, :7


def _ld(m, name, axis):
    return np.longdouble(m[f"{name}_{axis}_hi"]) + np.longdouble(m[f"{name}_{axis}_lo"])


def momentum(m, rings):
    """The four momenta of test_conservation.cpp:52-89 (translational and
    angular, with and without the ring terms), in extended precision."""
    out = {k: np.zeros(3, np.longdouble) for k in ("t_noring", "a_noring", "t", "a")}
    for i in range(len(m)):
        gm = np.longdouble(m["GM"][i])
        v = np.array([_ld(m[i:i + 1], "v", c)[0] for c in "xyz"])
        r = np.array([_ld(m[i:i + 1], "r", c)[0] for c in "xyz"])
        gl = np.array([_ld(m[i:i + 1], "GL", c)[0] for c in "xyz"])
        dm = gm * v
        da = gm * np.cross(r, v)
        das = gl * gm
        out["t_noring"] += dm
        out["t"] += dm
        out["a_noring"] += da + das
        out["a"] += da + das
        if m["ring"][i] >= 0:
            rg = rings[m["ring"][i]]
            f = np.longdouble(rg["GM_ratio"]) / (1 - np.longdouble(rg["GM_ratio"]))
            out["t"] += dm * f
            out["a"] += da * f
            out["a"] += gl / np.sqrt((gl * gl).sum()) * (np.longdouble(rg["GL_R2"]) * gm * np.longdouble(m["R2"][i]))
    return out


def conservation_errors(history):
    """max over the run of (max - min) / max, as test_conservation.cpp:94-112."""
    hmax = {k: np.max([h[k] for h in history], axis=0) for k in history[0]}
    hmin = {k: np.min([h[k] for h in history], axis=0) for k in history[0]}
    d = {k: hmax[k] - hmin[k] for k in hmax}
    n = np.linalg.norm
    err = max(n(d["t"]) / n(hmax["t"]), n(d["a"]) / n(hmax["a"]))
    err_noring = max(n(d["t_noring"]) / n(hmax["t"]), n(d["a_noring"]) / n(hmax["a"]))
    return float(err), float(err / err_noring)


def oracle_subsystem():
    full = po.OSystem.load(CONFIGS["solar504"], pool_at_load=False)
    sub = full.subset([full.mid(s) for s in SIDS])
    sub.accel(0, pool=True)
    return full, sub


def adjust(s, saturn, ring_ratio):
    """the push adjustment of test_conservation.cpp:16-22"""
    s.scale_geopotential(saturn, 3)
    s.scale_ring(saturn, 1e7)
    s.scale(saturn, 1 / (1 + ring_ratio()))


def run_oracle():
    full, sub = oracle_subsystem()
    saturn = sub.mid("699")
    adjust(sub, saturn, lambda: sub.models()[1][sub.masses["ring"][saturn]]["GM_ratio"])
    sub.accel(0, pool=True)
    rings = sub.models()[1]
    hist = []
    for _ in range(STEPS):
        sub.integrate(900.0, 1, pool=True)
        hist.append(momentum(sub.masses, rings))
    return sub.masses.copy(), hist


def product_subsystem():
    full, sub = oracle_subsystem()
    gps, rings = sub.models()
    s = MSystem.from_masses(sub.masses.copy(), gps, rings)
    return s


def run_product(use_gpu):
    s = product_subsystem()
    saturn = s.get_mid("699")
    adjust(s, saturn, lambda: s.models()[1][s.masses["ring"][saturn]]["GM_ratio"])
    s.accel(0)
    rings = s.models()[1]
    hist = []
    for _ in range(STEPS):
        s.integrate(900.0, 1, use_gpu=use_gpu)
        hist.append(momentum(s.masses, rings))
    return s.masses, hist


@pytest.fixture(scope="module")
def oracle_run():
    return run_oracle()


def test_oracle_conserves_momentum(oracle_run):
    _, hist = oracle_run
    err, ring_factor = conservation_errors(hist)
    print(f"oracle: max relative momentum variation {err:.3e}, ring correction factor {ring_factor:.3e}")
    assert err < EPS_CONSERVATION
    assert ring_factor < EPS_RING


def test_cpu_backend_conserves_momentum_and_matches_oracle(oracle_run):
    m_o, _ = oracle_run
    m, hist = run_product(0)
    err, ring_factor = conservation_errors(hist)
    print(f"product CPU: max relative momentum variation {err:.3e}, ring correction factor {ring_factor:.3e}")
    assert err < EPS_CONSERVATION
    assert ring_factor < EPS_RING
    for f in ("r_x_hi", "r_x_lo", "v_y_hi", "GL_z_hi", "s_xz_hi", "w_x", "GM"):
        assert np.array_equal(m[f].view(np.uint64), m_o[f].view(np.uint64)), f


def test_spin_invariants_of_an_isolated_body():
    """A lone body with spin and harmonics: |GL| stays constant (no torque)
    and the frame s stays orthonormal, over 2,000 CPU RK12 steps."""
    full, sub = oracle_subsystem()
    gps, rings = sub.models()
    one = sub.masses[:1].copy()
    one["ring"] = -1
    s = MSystem.from_masses(one, gps, rings)
    s.accel(0)
    gl0 = np.sqrt(sum(_ld(s.masses, "GL", c) ** 2 for c in "xyz"))
    for _ in range(20):
        s.integrate(900.0, 100, use_gpu=0)
    m = s.masses
    gl = np.sqrt(sum(_ld(m, "GL", c) ** 2 for c in "xyz"))
    assert abs(float(gl[0] / gl0[0] - 1)) < 1e-15
    S = np.array([[m[f"s_{a}{c}_hi"][0] + m[f"s_{a}{c}_lo"][0] for c in "xyz"] for a in "xyz"])
    assert np.max(np.abs(S @ S.T - np.eye(3))) < 1e-15


@pytest.mark.gpu
def test_gpu_conserves_momentum(oracle_run):
    print("\n[case] conservation: Saturn and four moons, 6000 GPU RK12 steps of 900 s", flush=True)
    m_o, _ = oracle_run
    m, hist = run_product(1)
    err, ring_factor = conservation_errors(hist)
    r = np.stack([m[f"r_{c}_hi"] + m[f"r_{c}_lo"] for c in "xyz"], -1)
    ro = np.stack([m_o[f"r_{c}_hi"] + m_o[f"r_{c}_lo"] for c in "xyz"], -1)
    dr = np.max(np.linalg.norm(r - ro, axis=1) / np.linalg.norm(ro, axis=1))
    print(f"GPU: max relative momentum variation {err:.3e}, ring correction factor {ring_factor:.3e}, "
          f"max rel dr vs oracle after {STEPS} steps {dr:.3e}", flush=True)
    assert err < EPS_CONSERVATION
    assert ring_factor < EPS_RING
    assert dr < 1e-12
