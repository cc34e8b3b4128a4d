"""Shared helpers for the tests."""
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
CONFIGS = {"small25": GOLDEN / "small25" / "Small_Config.txt",
           "full504": GOLDEN / "full504" / "Full_Config.txt",
           "solar504": GOLDEN / "solarsystem" / "SolarSystem_Config.txt"}


def load_npz(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.names:
        return all(bits_equal(a[f], b[f]) for f in a.dtype.names)
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def dd_value(m, name, axis):
    return m[f"{name}_{axis}_hi"] + m[f"{name}_{axis}_lo"]


def position(m):
    return np.stack([dd_value(m, "r", c) for c in "xyz"], -1)


def velocity(m):
    return np.stack([dd_value(m, "v", c) for c in "xyz"], -1)


def spin(m):
    return np.stack([dd_value(m, "GL", c) for c in "xyz"], -1)


def momenta(m, rings=None):
    """Linear and angular momentum (per G) of a system, as the reference's
    conservation test forms them (src/tests/test_conservation.cpp:55-80):
    sum GM v and sum GM (r x v + GL); with `rings`, a ring host also carries
    its ring's share of GM and the ring's own angular momentum."""
    gm = m["GM"][:, None]
    r, v, gl = position(m), velocity(m), spin(m)
    p = gm * v
    a = gm * np.cross(r, v)
    l = a + gm * gl
    if rings is not None:
        for i in np.nonzero(m["ring"] >= 0)[0]:
            rg = rings[m["ring"][i]]
            f = rg["GM_ratio"] / (1 - rg["GM_ratio"])
            p[i] = p[i] + p[i] * f
            u = gl[i] / np.linalg.norm(gl[i])
            l[i] = l[i] + a[i] * f + u * (rg["GL_R2"] * m["GM"][i] * m["R2"][i])
    return {"linear": p.sum(0), "angular": l.sum(0)}
