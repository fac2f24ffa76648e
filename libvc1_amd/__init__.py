"""libvc1_amd — MI355X (gfx950) motion-estimation backend for the libvc1 VC-1 encoder.

The product is the C-ABI shared library ``libvc1_amd/libvc1me.so`` (declared in
``include/vc1me.h``) built from ``libvc1_amd/csrc``.  This package holds its
build script, a ctypes binding used by tests and benchmarks, and the
synthetic-sequence generator of the measured workloads.
"""
from .backend import Backend, BackendError, Frame, load_library  # noqa: F401

__all__ = ["Backend", "BackendError", "Frame", "load_library"]
