"""MI355X (gfx950) back end for Ephemeris' force evaluation and RK12 stage
engine.  The compute path is the HIP library ``lib/libephemeris_hip.so``
(C ABI in include/ephemeris_hip.h); this package binds it and provides the
input-format generator and host-side helpers."""
from .layout import MASS_DTYPE, GP_DTYPE, RING_DTYPE  # noqa: F401

__all__ = ["MASS_DTYPE", "GP_DTYPE", "RING_DTYPE"]
