"""MI355X-native (gfx950) ShEPhERD scoring and multi-restart alignment.

Host side: Python on PyTorch-ROCm (device memory and streams).  Compute: hand-written HIP
kernels in libfss_hip.so behind the C ABI of include/fss.h.  There is no CPU fallback: every
entry point raises if the HIP library is missing.
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401
from .constants import ALPHA, LAM_SCALING, P_TYPES  # noqa: F401
