"""MI355X (gfx950) Z4c evolution hot path: host mirror of the GR-Athena++ Z4c
class surface over the C-ABI of libz4c_mi355x.so."""
from ._capi import EXPORTS, Params, Z4cError, load  # noqa: F401
from .build import LIB as LIB_PATH  # noqa: F401


def __getattr__(name):
    # torch-dependent classes are imported lazily so that the C-ABI can be
    # inspected on hosts without a GPU.
    if name in ("Z4c", "MeshBlockGeom", "BT_RK4_A", "BT_RK4_B"):
        from . import z4c as _z
        return getattr(_z, name)
    raise AttributeError(name)
