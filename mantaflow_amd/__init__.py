"""mantaflow_amd — MI355X-native per-step grid path of mantaflow.

libmfx.so (HIP kernels for gfx950 behind a C ABI, include/mfx.h) does the
work; `mantaflow_amd.manta` is the reference's Python scene API on top of it
and loads libmfx.so on import (raising if it is missing).  The package itself
stays importable without the library so that `python -m mantaflow_amd.build`
can build it.
"""
__all__ = ["manta", "build"]
