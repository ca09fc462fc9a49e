"""lqt1_amd: MI355X (gfx950) TPORd monodomain solver behind the C ABI of
include/lqt.h (lqt1_amd/liblqt.so).

Python here is the host-side mirror used by tests and the benchmark: it
loads the library and calls it; all compute runs in the library's HIP
kernels.
"""
from ._lib import (EPI, ENDO, FP32, FP64, LVENDO, LVEPI, MCELL, NONE, LqtError, Params, check,
                   device_count, device_name, ghost_map, lib, preset, suprep_fractions)
from .cells import Cells, TPORd
from .strands import Strands, StrandSpec, transmural
from .tissue import Tissue

__all__ = [
    "EPI", "ENDO", "FP32", "FP64", "LVENDO", "LVEPI", "MCELL", "NONE", "LqtError", "Params", "check",
    "device_count", "device_name", "ghost_map", "lib", "preset", "suprep_fractions", "Cells", "TPORd",
    "Strands", "StrandSpec", "transmural", "Tissue",
]
