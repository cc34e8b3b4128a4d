"""Shared helpers for the tests."""
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
CONFIGS = {"small25": GOLDEN / "small25" / "Small_Config.txt",
           "full504": GOLDEN / "full504" / "Full_Config.txt"}


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
