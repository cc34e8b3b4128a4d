"""The product loader's epoch correction and its harmonics adjustment, on the
host (no GPU), after msystem::load (src/modules/mass_loader.cpp:67-294) and
msystem::load(initial, ...) (:324-561):

* a nonzero TDB (initial ephemeris time) re-bases the time-dependent
  parameters to t = 0 (mass_loader.cpp:231-237): GM0 = GM - dGM TDB,
  exJ2 = dJ2 TDB and the C_static diagonal shifted by exJ2 (-1/2, -1/2, +1),
  so that msystem::update(TDB) (mass.cpp:28-45) gives back the loaded GM
  instead of jumping by dGM TDB.  TDB is read as a real number and rejected
  unless it is an integer (:226-229).  Checked in closed form against the
  same system loaded at TDB 0, and bit for bit against the oracle's loader;
* the loader removes the initial deformation from the static harmonics
  (mass_loader.cpp:543-551): C_static becomes C_static - (local(C_potential)
  - C_static) and C_potential the world-frame input harmonics.  So a fresh
  force evaluation's deformation pass, on the unchanged state, must give back
  the C_potential the loader stored.  With the adjustment's sign turned round
  it would differ by twice the deformation, which is checked to be far above
  the tolerance."""
import shutil

import numpy as np
import pytest

from ephemeris_amd import _lib
from ephemeris_amd.layout import MASS_DTYPE
from ephemeris_amd.system import MSystem
from oracle import pyoracle as po
from tests.util import CONFIGS, bits_equal

TDB = 631152000                                   # 20 Julian years in seconds


def _config_with_tdb(tmp_path, tdb_text):
    d = tmp_path / "full504"
    shutil.copytree(CONFIGS["full504"].parent, d)
    cfg = d / CONFIGS["full504"].name
    lines = [("TDB\t" + tdb_text) if ln.startswith("TDB") else ln for ln in cfg.read_text().splitlines()]
    cfg.write_text("\n".join(lines) + "\n")
    return cfg


def _mat(m, name):
    return np.stack([np.stack([m[f"{name}_{a}{b}"] for b in "xyz"], -1) for a in "xyz"], -2)


def test_nonzero_tdb_rebases_time_dependent_parameters(tmp_path):
    m0 = MSystem(CONFIGS["full504"]).masses
    s = MSystem(_config_with_tdb(tmp_path, str(TDB)))
    m = s.masses
    assert s.t_eph() == (float(TDB), 0.0)
    assert np.count_nonzero(m["dGM"]) and np.count_nonzero(m["dJ2"])     # MassRate and J2Rate bodies
    # closed form, the same double operations as the loader
    assert bits_equal(m["GM0"], m["GM"] - m["dGM"] * float(TDB))
    assert bits_equal(m["exJ2"], m["dJ2"] * float(TDB))
    assert bits_equal(m["C_static_xx"], m0["C_static_xx"] - m["exJ2"] / 2)
    assert bits_equal(m["C_static_yy"], m0["C_static_yy"] - m["exJ2"] / 2)
    assert bits_equal(m["C_static_zz"], m0["C_static_zz"] + m["exJ2"])
    for f in ("GM", "r_x_hi", "v_y_lo", "C_potential_xz", "daccel_z", "C_static_xy"):   # nothing else moves
        assert bits_equal(m[f], m0[f]), f
    # and equal to the oracle's loader, every field the loader sets
    o = po.OSystem.load(_config_with_tdb(tmp_path / "o", str(TDB)), pool_at_load=False)
    ref = o.masses.copy()
    fields = [f for f in MASS_DTYPE.names if not f.startswith(("Egrad", "idaccel", "idtorque", "jdaccel", "jdtorque"))]
    assert not [f for f in fields if not bits_equal(m[f], ref[f])]


def test_update_at_tdb_gives_back_the_loaded_mass(tmp_path):
    s = MSystem(_config_with_tdb(tmp_path, str(TDB)))
    gm = s.masses["GM"].copy()
    s.update(float(TDB))
    after = s.masses
    moving = after["dGM"] != 0
    assert np.any(moving)
    # GM0 + dGM TDB = GM to within one rounding; without the correction it would be off by dGM TDB
    assert np.all(np.abs(after["GM"] - gm) <= 2 * np.spacing(np.abs(gm)))
    assert np.all(np.abs(after["dGM"][moving] * float(TDB)) > 1e3 * np.spacing(np.abs(gm[moving])))
    assert bits_equal(after["exJ2"], after["dJ2"] * float(TDB))


@pytest.mark.parametrize("text,ok", [("6.3e8", True), ("1.5", False), ("-86400", True), ("1e30", False)])
def test_tdb_must_be_an_integer(tmp_path, text, ok):
    cfg = _config_with_tdb(tmp_path, text)
    if ok:
        assert MSystem(cfg).t_eph() == (float(text), 0.0)
    else:
        with pytest.raises(_lib.EphError, match="integer"):
            MSystem(cfg)


@pytest.mark.parametrize("name", ["full504", "solar504"])
def test_fresh_deformation_gives_back_the_loaded_harmonics(name):
    s = MSystem(CONFIGS[name], threads=4)
    loaded = s.masses
    c0 = _mat(loaded, "C_potential")
    s.accel(-1)                                   # deformation pass (and force) on the unchanged state
    c1 = _mat(s.masses, "C_potential")
    scale = np.abs(c0).max(axis=(1, 2))
    deformed = scale > 0
    err = np.abs(c1 - c0).max(axis=(1, 2))[deformed] / scale[deformed]
    assert err.max() < 1e-13, err.max()
    # the deformation the loader removed: C_potential and C_static differ by it
    # (eigenvalues, since the two are in different frames)
    cs = _mat(loaded, "C_static")
    ev = np.linalg.eigvalsh(c0[deformed]) - np.linalg.eigvalsh(cs[deformed])
    rel = np.abs(ev).max(axis=1) / scale[deformed]
    assert rel.max() > 1e-6, rel.max()
