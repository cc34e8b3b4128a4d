"""Byte layout of the boundary structs against the reference's class mass,
and the binary checkpoint format (src/modules/mass_loader.cpp:566-759).

The offset check compiles a translation unit that includes the reference's
src/physics/mass.h beside include/ephemeris_hip.h and static_asserts that
every eph_mass member sits at the offset of its class mass counterpart over
the bytes the checkpoint stores (r .. C_potential, mass.h:67-131); it is
skipped where the reference checkout is absent."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ephemeris_amd.layout import MASS_DTYPE
from ephemeris_amd.system import MSystem
from tests.util import CONFIGS, GOLDEN, ROOT, bits_equal

REF = Path("/root/reference/src")
MEMBERS = ["r", "v", "GL", "s", "w", "GM", "exJ2", "sid", "GM0", "dGM", "dJ2", "R", "inertia", "k2", "k2r",
           "tide_delay", "tide_delay_factor", "lum", "recpt", "A", "R2", "sR2", "rR2G_4c", "C_static",
           ("gp", "gpmodel"), ("ring", "ringmodel"), "phi", "beta2", "naccel", "daccel", "dtorque", "gaccel",
           "beta", "GI", "C_potential", "Egrad"]


@pytest.mark.skipif(not (REF / "physics" / "mass.h").exists(), reason="reference checkout not present")
def test_eph_mass_offsets_equal_class_mass(tmp_path):
    lines = ['#include "physics/mass.h"', '#include "ephemeris_hip.h"', "#include <cstddef>"]
    for m in MEMBERS:
        a, b = (m, m) if isinstance(m, str) else m
        lines.append(f'static_assert(offsetof(eph_mass, {a}) == offsetof(mass, {b}), "{a}");')
    lines.append('static_assert(sizeof(real) == 16 && sizeof(mpvec) == 48 && sizeof(fast_mpmat) == 72, "sizes");')
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines) + "\n")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-w", "-include", "cstddef", "-include", "atomic", "-include", "thread",
           "-include", "cstring", f"-I{REF}", f"-I{ROOT / 'include'}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def _checkpoint(s, path):
    s.save_checkpoint(path)
    return Path(path).read_bytes()


def test_checkpoint_layout(tmp_path):
    s = MSystem(CONFIGS["solar504"])
    raw = _checkpoint(s, tmp_path / "c.dat")
    magic, version, t_hi, t_lo, dt, integ = struct.unpack_from("<qqdddq", raw, 0)
    cad, maxlen, cdt, gmc, gmp, gmt, pmc, nmass, nbary, msize = struct.unpack_from("<dddddddqqq", raw, 48)
    assert magic == 0x53484c486d687045 and version == 4          # src/configs.cpp:5-6
    assert (t_hi, t_lo, dt, integ) == (0.0, 0.0, 300.0, 2)
    assert (cad, maxlen, cdt, gmc, gmp, gmt, pmc) == (7200.0, 631152000.0, 28800.0, 1e13, 2e17, 14e11, 0.0)
    assert nmass == 504 and nbary == len(s.barycens()) and msize == 824
    off = 128
    for pid, hid, gid, tid, mid, ch, _ in s.barycens():           # barycen_ids + children
        assert struct.unpack_from("<qqqqqq", raw, off) == (pid, hid, gid, tid, mid, len(ch))
        off += 48 + 8 * len(ch)
    m = s.masses
    gps, rings = s.models()
    for i in range(nmass):
        rec = np.frombuffer(raw[off:off + 824] + bytes(MASS_DTYPE.itemsize - 824), MASS_DTYPE)[0]
        for f in ("r_x_hi", "v_z_lo", "GL_y_hi", "s_zz_lo", "GM", "sid", "C_static_zz", "C_potential_xy", "beta_z"):
            assert bits_equal(np.array(rec[f]), np.array(m[f][i])), (i, f)
        off += 824
        if m["gp"][i] >= 0:
            g = gps[m["gp"][i]]
            size = 16 + 8 * (g["Nz"] * (g["Nz"] + 2) - 3)
            assert rec["gp"] == size                                 # the model pointer holds its size
            assert struct.unpack_from("<qq", raw, off) == (g["Nz"], g["Nt"])
            off += size
        else:
            assert rec["gp"] == 0
        if m["ring"][i] >= 0:
            rg = rings[m["ring"][i]]
            size = 40 + 24 * rg["N"]
            assert rec["ring"] == size
            n, ratio = struct.unpack_from("<qd", raw, off)
            assert n == rg["N"] and ratio == rg["GM_ratio"]
            disk0 = struct.unpack_from("<ddd", raw, off + 40)     # {GsR2, R_R, H_R} of disk 0
            assert disk0 == (rg["GsR2"][0], rg["R_R"][0], rg["H_R"][0])
            off += size
        else:
            assert rec["ring"] == 0
    assert off == len(raw)


def test_checkpoint_round_trip_resumes_bit_identically(tmp_path):
    """save -> load -> integrate equals an uninterrupted run, bit for bit
    (CPU back end), for the reference's Solar System."""
    a = MSystem(CONFIGS["solar504"], threads=8)
    a.integrate(300.0, 2, use_gpu=0)
    a.save_checkpoint(tmp_path / "mid.dat")
    b = MSystem.from_checkpoint(tmp_path / "mid.dat", threads=8)
    ma, mb = a.masses, b.masses
    prefix = MASS_DTYPE.names[:MASS_DTYPE.names.index("Egrad_x")]
    assert all(bits_equal(ma[f], mb[f]) for f in prefix if f not in ("gp", "ring"))
    assert b.t_eph() == a.t_eph() and b.params() == a.params()
    assert [x[:6] for x in b.barycens()] == [x[:6] for x in a.barycens()]
    ga, ra = a.models()
    gb, rb = b.models()
    assert bits_equal(ga, gb) and bits_equal(ra, rb)
    a.integrate(300.0, 2, use_gpu=0)
    b.integrate(300.0, 2, use_gpu=0)
    ma, mb = a.masses, b.masses
    assert all(bits_equal(ma[f], mb[f]) for f in prefix), [f for f in prefix if not bits_equal(ma[f], mb[f])]


def test_corrupt_checkpoints_are_rejected(tmp_path):
    from ephemeris_amd import _lib
    s = MSystem(CONFIGS["small25"])
    raw = _checkpoint(s, tmp_path / "c.dat")
    for bad in (raw[:100], raw[:len(raw) // 2], b"x" * 8 + raw[8:], raw[:8] + struct.pack("<q", 3) + raw[16:]):
        p = tmp_path / "bad.dat"
        p.write_bytes(bad)
        with pytest.raises(_lib.EphError):
            MSystem.from_checkpoint(p)
