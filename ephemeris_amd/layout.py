"""numpy views of the C-ABI structs of include/ephemeris_hip.h.

``MASS_DTYPE`` is eph_mass: a field-for-field mirror of class mass
(src/physics/mass.h:67-143) with every member 8 bytes wide; double-double
members appear as ``<name>_<axis>_hi`` / ``_lo``."""
import numpy as np


def _v(name):
    return [(name + "_" + c, "f8") for c in "xyz"]


def _ddv(name):
    return [(f"{name}_{c}_{p}", "f8") for c in "xyz" for p in ("hi", "lo")]


def _m(name):
    return [(f"{name}_{a}{c}", "f8") for a in "xyz" for c in "xyz"]


def _ddm(name):
    return [(f"{name}_{a}{c}_{p}", "f8") for a in "xyz" for c in "xyz" for p in ("hi", "lo")]


MASS_DTYPE = np.dtype(
    _ddv("r") + _ddv("v") + _ddv("GL") + _ddm("s") + _v("w")
    + [("GM", "f8"), ("exJ2", "f8"), ("sid", "u8")]
    + [(k, "f8") for k in ("GM0", "dGM", "dJ2", "R", "inertia", "k2", "k2r", "tide_delay",
                           "tide_delay_factor", "lum", "recpt", "A", "R2", "sR2", "rR2G_4c")]
    + _m("C_static") + [("gp", "i8"), ("ring", "i8"), ("phi", "f8"), ("beta2", "f8")]
    + _v("naccel") + _v("daccel") + _v("dtorque") + _v("gaccel") + _v("beta")
    + _m("GI") + _m("C_potential")
    + _v("Egrad") + _v("idaccel") + _v("idtorque") + _v("jdaccel") + _v("jdtorque")
    + [("min_distance", "f8"), ("max_influence", "f8")])

GP_DTYPE = np.dtype([("Nz", "i8"), ("Nt", "i8"), ("c", "f8", (77,))])
RING_DTYPE = np.dtype([("N", "i8"), ("GM_ratio", "f8"), ("A", "f8"), ("J2", "f8"),
                       ("GL_R2", "f8"), ("GsR2", "f8", (64,)), ("R_R", "f8", (64,)),
                       ("H_R", "f8", (64,))])

# outputs of one force evaluation (class mass auxiliary variables, plus the
# spin w that the deform fixed point re-solves and the running extrema)
ACCEL_OUTPUTS = (["phi", "beta2"]
                 + [n for n, _ in _v("naccel") + _v("daccel") + _v("dtorque") + _v("gaccel")
                    + _v("beta") + _m("GI") + _m("C_potential") + _v("w")]
                 + ["min_distance", "max_influence"])


def field_names(prefix):
    return [n for n in MASS_DTYPE.names if n == prefix or n.startswith(prefix + "_")]
