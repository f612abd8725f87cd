"""mantaflow_amd — MI355X-native per-step grid path of mantaflow.

libmfx.so (HIP kernels for gfx950 behind a C ABI, include/mfx.h) does the
work; `mantaflow_amd.manta` is the reference's Python scene API on top of it.
Importing this package loads libmfx.so and fails if it is missing.
"""
from . import _abi

_abi.lib()

__all__ = ["manta"]
