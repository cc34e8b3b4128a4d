"""The product loader (text inputs -> resident system through the C ABI)
against the oracle's restatement of msystem::load (mass_loader.cpp:324-561),
and the msystem-style entry points (accel / integrate / diagnostics)."""
import numpy as np
import pytest

from ephemeris_amd import _lib
from ephemeris_amd.layout import MASS_DTYPE
from ephemeris_amd.system import MSystem
from oracle import pyoracle as po
from tests.util import CONFIGS, bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["small25", "full504"])
def test_load_matches_oracle(name):
    print(f"\n[case] load {name}", flush=True)
    m = MSystem(CONFIGS[name]).masses
    o = po.OSystem.load(CONFIGS[name], pool_at_load=False)
    ref = o.masses
    ring = ref["ring"] >= 0
    bad = {f: int((m[f].view(np.uint64) != ref[f].view(np.uint64)).sum())
           for f in MASS_DTYPE.names if f not in ("Egrad_x", "Egrad_y", "Egrad_z") and not bits_equal(m[f], ref[f])}
    print("load mismatches:", bad, flush=True)
    if not ring.any():
        assert not bad
    else:
        for f in bad:
            a, b = m[f].astype(float), ref[f].astype(float)
            rel = np.max(np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-300))
            assert rel < 1e-13, (f, rel)


def test_cpu_backend_is_not_provided():
    s = MSystem(CONFIGS["small25"])
    with pytest.raises(_lib.EphError) as e:
        s.integrate(300.0, 1, use_gpu=0)
    assert e.value.code == _lib.EPH_ENOTSUP


def test_integrate_backwards_returns_close_to_start():
    s = MSystem(CONFIGS["small25"])
    r0 = s.masses["r_x_hi"].copy()
    s.integrate(300.0, 24)
    s.integrate(300.0, -24)
    r1 = s.masses["r_x_hi"]
    assert np.max(np.abs(r1 - r0) / np.abs(r0)) < 1e-12
    assert s.t_eph()[0] + s.t_eph()[1] == 0.0


def test_timestep_diagnostic_reported():
    s = MSystem(CONFIGS["small25"])
    s.integrate(40000.0, 1)          # far too large for the close moons
    d = s.diagnostics()
    assert any(k == "timestep" for _, k, _, _ in d)
