"""The oracle (C restatement) against golden vectors produced by the
reference's own code (tests/golden/make_golden.py), bit for bit."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.util import GOLDEN, bits_equal, load_npz

FN = None


def fn():
    global FN
    if FN is None:
        FN = load_npz(GOLDEN / "functions.npz")
    return FN


def test_dd_ops_and_sqrt():
    f = fn()
    L = po.lib()
    out = np.empty(8)
    sq = np.empty(2)
    for i in range(len(f["dd_a"])):
        L.o_dd_ops(po.dptr(np.ascontiguousarray(f["dd_a"][i])), po.dptr(np.ascontiguousarray(f["dd_b"][i])), po.dptr(out))
        assert bits_equal(out, f["dd_ops"][i]), i
        L.o_dd_sqrt(po.dptr(np.ascontiguousarray(f["dd_sqrt_in"][i])), po.dptr(sq))
        assert bits_equal(sq, f["dd_sqrt"][i]), i


def test_rotation_inverse_normalize_ascnode():
    f = fn()
    L = po.lib()
    o9 = np.empty(9)
    o3 = np.empty(3)
    for i in range(len(f["rot_w"])):
        L.o_rotation_matrix(po.dptr(np.ascontiguousarray(f["rot_w"][i])), float(f["rot_t"][i]), po.dptr(o9))
        assert bits_equal(o9, f["rot"][i]), ("rotation", i)
        L.o_inverse(po.dptr(np.ascontiguousarray(f["inv_in"][i])), po.dptr(o9))
        assert bits_equal(o9, f["inv"][i]), ("inverse", i)
        v = np.ascontiguousarray(f["norm_in"][i].copy())
        n = L.o_normalize(po.dptr(v))
        assert bits_equal(np.array([n]), f["norm"][i:i + 1]) and bits_equal(v, f["norm_vec"][i]), ("normalize", i)
        L.o_asc_node(po.dptr(np.ascontiguousarray(f["rot_w"][i])), po.dptr(o3))
        assert bits_equal(o3, f["asc"][i]), ("asc_node", i)


def test_geopotential_sum():
    f = fn()
    gp = np.ascontiguousarray(f["gp"])
    o = np.empty(3)
    for i, r in enumerate(f["gp_r"]):
        po.lib().o_gp_sum(gp.ctypes.data, float(f["gp_R"]), po.dptr(np.ascontiguousarray(r)), po.dptr(o))
        assert bits_equal(o, f["gp_out"][i]), i


@pytest.mark.parametrize("which", ["ring", "ring_thick"])
def test_ring_sum(which):
    f = fn()
    rg = np.ascontiguousarray(f[which])
    o = np.empty(3)
    for i, r in enumerate(f["ring_r"]):
        po.lib().o_ring_sum(rg.ctypes.data, float(f["ring_R"]), po.dptr(np.ascontiguousarray(r)), po.dptr(o))
        assert bits_equal(o, f[which + "_out"][i]), i


CONFIG = {"small25": GOLDEN / "small25" / "Small_Config.txt", "full504": GOLDEN / "full504" / "Full_Config.txt",
          "solar504": GOLDEN / "solarsystem" / "SolarSystem_Config.txt"}


@pytest.mark.parametrize("name", ["small25", "full504", "solar504"])
def test_deform_by(name):
    d = load_npz(GOLDEN / f"deform_{name}.npz")
    s = po.OSystem.load(CONFIG[name])
    for k, i in enumerate(d["idx"]):
        m = s.masses
        m[:] = d["masses_in"]
        po.lib().o_deform_by(po.lib().o_masses(s.h), len(s), int(i))
        for fld in ("phi", "naccel_x", "naccel_y", "naccel_z") + tuple(n for n in po.MASS_DTYPE.names if n.startswith(("C_potential", "GI_"))):
            assert bits_equal(np.array([m[fld][i]]), np.array([d["out"][fld][k]])), (i, fld)


@pytest.mark.parametrize("name,tag", [("small25", "serial"), ("full504", "pool"), ("full504", "serial"),
                                      ("solar504", "pool"), ("solar504", "serial")])
def test_accel_matches_reference(name, tag):
    d = load_npz(GOLDEN / f"accel_{name}_{tag}.npz")
    s = po.OSystem.load(CONFIG[name])
    m = s.masses
    m[:] = d["masses_in"]
    s.accel(0 if tag == "pool" else -1, pool=True)
    bad = [f for f in po.MASS_DTYPE.names if not bits_equal(m[f], d["masses_out"][f])]
    assert not bad, bad
