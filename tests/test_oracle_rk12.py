"""The RK12 stage engine of the oracle, pinned by mathematics: the tableau's
order conditions in exact rational arithmetic, and the measured convergence
order of a Sun/planet/moon integration (design of the reference's
src/tests/test_order.cpp:35-101: order > 11.375 for r and v, > 2.5 for the
spin w and the body axes)."""
import math
import tempfile
from fractions import Fraction as F

import numpy as np
import pytest

from ephemeris_amd import generator as g
from oracle import pyoracle as po


def tableau():
    T = po.rk12_table()
    a = {}
    for k in range(1, 26):
        for j in range(k):
            hi, lo = T[k * (k - 1) // 2 + j]
            a[(k, j)] = F(hi) + F(lo)
    c = [F(0)] + [sum(a[(k, j)] for j in range(k)) for k in range(1, 25)]
    b = [a[(25, j)] for j in range(25)]
    return a, b, c


def test_quadrature_conditions():
    a, b, c = tableau()
    for q in range(12):
        res = sum(b[j] * c[j] ** q for j in range(25)) - F(1, q + 1)
        assert abs(float(res)) < 1e-30, (q, float(res))


def test_bac_conditions():
    a, b, c = tableau()
    for q in range(10):
        res = sum(b[j] * sum(a[(j, k)] * c[k] ** q for k in range(j)) for j in range(1, 25)) - F(1, (q + 1) * (q + 2))
        assert abs(float(res)) < 1e-30, (q, float(res))


def test_row_sums_in_unit_interval_and_sparse():
    a, b, c = tableau()
    nz = sum(1 for v in a.values() if v != 0)
    assert nz == 191
    assert all(0 <= float(x) <= 1 for x in c)


def _triple():
    rng = np.random.default_rng(7)
    star = g._star(rng, tides=False)
    mass = 6.4e23
    r, v, nrm = g.kepler_state(star.GM + g.G * mass, 1.52 * g.AU, 0.09, 0.03, 0.5, 1.0, 2.0)
    x, z, w = g._spin_frame(rng, nrm, 24.6, 0.4)
    p = g.Body("Planet", "P1", g.G * mass, 3390.0, inertia=0.366, k2=0.17, k2r=0.17, td=600.0, J2=1.96e-3,
               r=r, v=v, x_axis=x, z_axis=z, w=w)
    mm, am = 1.07e16, 9.376e6
    rm, vm, nm = g.kepler_state(p.GM + g.G * mm, am, 0.015, 0.02, 0.3, 0.2, 0.1)
    per = 2 * math.pi * math.sqrt(am ** 3 / (p.GM + g.G * mm)) / 3600
    xm, zm, wm = g._spin_frame(rng, nm, per, 0.0)
    m = g.Body("Moon", "M1", g.G * mm, 11.1, inertia=0.4, k2=0.01, k2r=0.01, td=100.0, J2=0.1, C22=0.01,
               r=p.r + rm, v=p.v + vm, x_axis=xm, z_axis=zm, w=wm)
    bodies = [star, p, m]
    g._barycentre(bodies)
    return g.SystemSpec(bodies)


def _order_solve(dratio, dt0, dt1, dt2):            # test_order.cpp:9-33
    b0, b1 = dt2 / dt1, dt0 / dt1
    lb0, lb1 = math.log(b0), math.log(b1)
    order = max(math.log(dratio) / (lb0 - lb1) * 2, 3)
    best, last, n = float("nan"), float("inf"), 0
    while True:
        b1o, b0o = b1 ** order, b0 ** order
        diff = dratio * (b1o - 1) - (1 - b0o)
        if not abs(diff) < last:
            break
        best, last = order, abs(diff)
        order -= diff / (b0o * lb0 + b1o * dratio * lb1)
        n += 1
    return best if n >= 2 else float("nan")


def test_convergence_order():
    d = tempfile.mkdtemp()
    base = po.OSystem.load(g.write_system(_triple(), d, "Tri"))
    dts, lcm = [1200, 1080, 960, 880, 800, 720, 640, 600], 950400
    states = []
    for dt in dts:
        s = base.clone()
        s.integrate(dt, lcm // dt)
        m = s.masses
        rel = lambda f: np.array([m[f"{f}_{c}_hi"][2] + m[f"{f}_{c}_lo"][2] - m[f"{f}_{c}_hi"][1] - m[f"{f}_{c}_lo"][1] for c in "xyz"])
        axis = lambda a: np.array([m[f"s_{a}{c}_hi"][2] + m[f"s_{a}{c}_lo"][2] for c in "xyz"])
        states.append([rel("r"), rel("v"), np.array([m[f"w_{c}"][2] for c in "xyz"]), axis("x"), axis("z")])
    for i in range(1, len(dts) - 1):
        for j in range(5):
            dr = np.linalg.norm(states[i][j] - states[i + 1][j]) / np.linalg.norm(states[i - 1][j] - states[i][j])
            assert 0 < dr < 1
            order = _order_solve(dr, dts[i - 1], dts[i], dts[i + 1])
            assert order > (11.375 if j < 2 else 2.5), ("rvwxz"[j], i, order)
