"""nemo_amd: MI355X-native (gfx950) spatial operator and RK stage update for
nemo, behind the C-ABI of include/nemo_gpu.h.

The compute path is the in-tree HIP library ``libnemo_gpu.so``; this package
is its host-side mirror (mesh upload format, initial data, the reference's
kernel/timedisc/driver call sequence).  Importing a compute entry point
without the built library raises: there is no CPU fallback.
"""
from ._lib import (  # noqa: F401
    Config, NemoGpuError, EQ_EULER, VOL_WEAK, VOL_CENTRAL, VOL_EC, SURF_RUSANOV, SURF_HLLC,
    IND_NONE, IND_PP, TD_LSRK54, TD_SSPRK54, LIB_PATH,
)

__all__ = ["Config", "NemoGpuError", "LIB_PATH"]
