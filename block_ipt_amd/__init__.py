"""MI355X-native block-IPT eigensolver.

The compute path is libipt_hip.so (hand-written HIP kernels for gfx950
behind the C ABI in include/ipt_hip.h).  This package only loads that
library and mirrors the reference's entry points for Python callers; it
raises immediately when the library is missing.
"""
from .abi import IPTCudaResult, IptError, lib, library_path  # noqa: F401
from . import api  # noqa: F401
