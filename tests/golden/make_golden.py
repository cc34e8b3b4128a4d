"""Golden vectors from the reference's own code (run here, where the
reference checkout exists; the fixtures travel, the reference does not).

The vectors come from oracle/_ref/libephref.so, built by `make -C oracle ref`
from the reference sources in place: the task bodies of src/physics/mass.cpp
(deform, main force, non-point rounds), geopotential.cpp, ring.cpp, the
force macros of mass.impl and the math headers.  Inputs are seeded; every
fixture stores inputs and reference outputs as float64/structured arrays
(np.savez, no pickle).

    python tests/golden/make_golden.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import pyoracle as po  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "libephref.so"


def ref_lib():
    if not REF.exists():
        import subprocess
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    L = C.CDLL(str(REF))
    P, D, I64, I = C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_int
    L.ref_accel.argtypes = [P, I64, P, I64, P, I64, I]
    L.ref_accel.restype = I
    L.ref_deform_by.argtypes = [P, I64, I64, P, I64, P, I64]
    L.ref_gp_sum.argtypes = [P, C.c_double, D, D]
    L.ref_ring_sum.argtypes = [P, C.c_double, D, D]
    L.ref_dd_ops.argtypes = [D, D, D]
    L.ref_dd_sqrt.argtypes = [D, D]
    L.ref_rotation_matrix.argtypes = [D, C.c_double, D]
    L.ref_inverse.argtypes = [D, D]
    L.ref_normalize.argtypes = [D]
    L.ref_normalize.restype = C.c_double
    L.ref_asc_node.argtypes = [D, D]
    L.ref_analyse.restype = I64
    L.ref_analyse.argtypes = [P, I64, C.c_double, C.c_double, P, I64, P, I64, P]
    return L


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


CASES = {"small25": (HERE / "small25" / "Small_Config.txt", [1]),
         "full504": (HERE / "full504" / "Full_Config.txt", [0, 1]),
         # the reference's own Solar System (initials/SolarSystem, copied as
         # data): tesseral geopotentials of the Moon and Mars, Saturn's ring
         "solar504": (HERE / "solarsystem" / "SolarSystem_Config.txt", [0, 1])}


def accel_fixtures(L, names=("small25", "full504")):
    for name in names:
        cfg, modes = CASES[name]
        s = po.OSystem.load(cfg, pool_at_load=False)
        gps, rings = s.models()
        for mode in modes:
            inp = s.masses.copy()
            out = np.ascontiguousarray(inp.copy())
            rc = L.ref_accel(out.ctypes.data, len(out), gps.ctypes.data if len(gps) else None, len(gps),
                             rings.ctypes.data if len(rings) else None, len(rings), mode)
            assert rc == 0
            tag = "pool" if mode == 0 else "serial"
            np.savez_compressed(HERE / f"accel_{name}_{tag}.npz", masses_in=inp, masses_out=out,
                                gps=gps, rings=rings)
        # deform_by of a few bodies (and of every body with a model)
        idx = np.arange(0, len(s), max(1, len(s) // 16))
        if name == "solar504":
            idx = np.union1d(idx, np.nonzero((s.masses["gp"] >= 0) | (s.masses["ring"] >= 0))[0])
        outs = []
        for i in idx:
            m = np.ascontiguousarray(s.masses.copy())
            L.ref_deform_by(m.ctypes.data, len(m), int(i), gps.ctypes.data if len(gps) else None, len(gps),
                            rings.ctypes.data if len(rings) else None, len(rings))
            outs.append(m[i])
        np.savez_compressed(HERE / f"deform_{name}.npz", masses_in=s.masses.copy(), idx=idx,
                            out=np.array(outs), gps=gps, rings=rings)


def function_fixtures(L):
    rng = np.random.default_rng(4242)
    out = {}
    # double-double + - * / and sqrt
    n = 2000
    a = np.empty((n, 2)); b = np.empty((n, 2))
    ah = rng.normal(size=n) * 10.0 ** rng.integers(-30, 30, n)
    bh = rng.normal(size=n) * 10.0 ** rng.integers(-30, 30, n)
    a[:, 0], a[:, 1] = ah, ah * rng.uniform(-1, 1, n) * 2.0 ** -53
    b[:, 0], b[:, 1] = bh, bh * rng.uniform(-1, 1, n) * 2.0 ** -53
    r = np.empty((n, 8)); sq = np.empty((n, 2))
    sq_in = np.where(a[:, :1] > 0, a, -a)
    for i in range(n):
        L.ref_dd_ops(dp(a[i]), dp(b[i]), dp(r[i]))
        L.ref_dd_sqrt(dp(np.ascontiguousarray(sq_in[i])), dp(sq[i]))
    out.update(dd_a=a, dd_b=b, dd_ops=r, dd_sqrt_in=sq_in, dd_sqrt=sq)
    # rotation_matrix, inverse, normalize, asc_node
    m = 500
    w = rng.normal(size=(m, 3)) * 10.0 ** rng.integers(-8, 0, (m, 1))
    w[:5] = [[0, 0, 0], [1e-300, 0, 0], [0, 0, 1e-170], [3e200, 1, 0], [0, 0, -7e-5]]
    t = rng.uniform(-1e4, 1e4, m)
    rot = np.empty((m, 9))
    for i in range(m):
        L.ref_rotation_matrix(dp(np.ascontiguousarray(w[i])), float(t[i]), dp(rot[i]))
    mats = rng.normal(size=(m, 9))
    inv = np.empty((m, 9))
    for i in range(m):
        L.ref_inverse(dp(np.ascontiguousarray(mats[i])), dp(inv[i]))
    vv = np.ascontiguousarray(w.copy())
    nrm = np.empty(m)
    for i in range(m):
        nrm[i] = L.ref_normalize(dp(vv[i]))
    asc = np.empty((m, 3))
    for i in range(m):
        L.ref_asc_node(dp(np.ascontiguousarray(w[i])), dp(asc[i]))
    out.update(rot_w=w, rot_t=t, rot=rot, inv_in=mats, inv=inv, norm_in=w, norm_vec=vv, norm=nrm, asc=asc)
    # geopotential, degrees 2..8, zonal and tesseral
    gp = np.zeros(1, po.GP_DTYPE)
    gp["Nz"], gp["Nt"] = 8, 8
    gp["c"][0] = rng.normal(size=77) * 1e-3
    R = 6.0e6
    pts = rng.normal(size=(600, 3))
    pts /= np.linalg.norm(pts, axis=1)[:, None]
    pts *= R * rng.uniform(1.0, 40.0, (600, 1))
    pts[:20, :2] *= 1e-160                       # near the poles
    pts[20:25, :2] = 0
    g = np.empty((len(pts), 3))
    for i in range(len(pts)):
        L.ref_gp_sum(gp.ctypes.data, R, dp(np.ascontiguousarray(pts[i])), dp(g[i]))
    out.update(gp=gp, gp_R=np.float64(R), gp_r=pts, gp_out=g)
    # ring: the full system's ring and a thick variant, points over all branches
    full = po.OSystem.load(HERE / "full504" / "Full_Config.txt")
    _, rings = full.models()
    ring = rings[:1].copy()
    thick = ring.copy()
    thick["H_R"][0, :int(thick["N"][0])] = 0.05
    host_R = 5.8e7
    rr = []
    for rho in np.concatenate([np.linspace(0.2, 3.0, 57), [1.0, 1.24, 1.53, 2.27, 2.271]]):
        for z in [0.0, 1e-9, 1e-4, 0.003, 0.05, 0.4, 1.5, -0.2, -3.0]:
            ph = rng.uniform(0, 2 * np.pi)
            rr.append([rho * np.cos(ph), rho * np.sin(ph), z])
    rr = np.array(rr) * host_R
    o1 = np.empty_like(rr); o2 = np.empty_like(rr)
    for i in range(len(rr)):
        L.ref_ring_sum(ring.ctypes.data, host_R, dp(np.ascontiguousarray(rr[i])), dp(o1[i]))
        L.ref_ring_sum(thick.ctypes.data, host_R, dp(np.ascontiguousarray(rr[i])), dp(o2[i]))
    out.update(ring=ring, ring_thick=thick, ring_R=np.float64(host_R), ring_r=rr, ring_out=o1, ring_thick_out=o2)
    np.savez_compressed(HERE / "functions.npz", **out)


def analyse_fixtures(L, names):
    """msystem::analyse(true) of the reference (barycen_analyzer.cpp) on the
    loaded bodies: the orbit hierarchy and each node's GM_sys."""
    for name in names:
        cfg, _ = CASES[name]
        s = po.OSystem.load(cfg, pool_at_load=False)
        m = np.ascontiguousarray(s.masses.copy())
        n = len(m)
        cap = 2 * n + 16
        ids = np.zeros((cap, 7), np.int64)
        ch = np.zeros(4 * cap, np.int64)
        gm = np.zeros((cap, 2))
        nb = L.ref_analyse(m.ctypes.data, n, 0.0, 0.0, ids.ctypes.data, cap, ch.ctypes.data, len(ch), gm.ctypes.data)
        assert 0 < nb <= cap and ids[:nb, 5].sum() <= len(ch)
        np.savez_compressed(HERE / f"analyse_{name}.npz", masses_in=m, ids=ids[:nb], gm_sys=gm[:nb],
                            children=ch[:ids[:nb, 5].sum()])


if __name__ == "__main__":
    L = ref_lib()
    if len(sys.argv) > 1 and sys.argv[1] == "solar":
        # only the fixtures of the reference's Solar System and the analyses
        accel_fixtures(L, ["solar504"])
        analyse_fixtures(L, ["small25", "solar504"])
    else:
        accel_fixtures(L)
        function_fixtures(L)
        analyse_fixtures(L, ["small25", "solar504"])
    print("fixtures written to", HERE)
