"""melt_ibm_amd: MI355X (gfx950) implementation of the melt_IBM fluid /
immersed-boundary RK3 substep behind the reference's subroutine interface.

The compute lives in libmeltgpu.so (hand-written HIP, C ABI in
include/melt_ibm.h); this package is the host-side mirror of the reference's
driver interface plus input generators.  Importing it does not load the
library; MeltContext does, and raises if the library has not been built.
"""
from .api import Body, MeltConfig, MeltContext, MeltError, device_available, load_library  # noqa: F401

__all__ = ["Body", "MeltConfig", "MeltContext", "MeltError", "device_available", "load_library"]
