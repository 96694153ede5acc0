"""MI355X-native matrix-free operator evaluation for PRISMS-PF applications.

The compute path is HIP (gfx950) behind the C-ABI of ``include/ppf.h``; one shared
library per application lives in ``phasefield_amd/lib``.  ``phasefield_amd.capi`` is
the ctypes binding, ``phasefield_amd.build`` compiles the libraries, and
``phasefield_amd.prm`` writes parameter files of the benchmark configurations.
"""

from .capi import Context, Library, PPFError  # noqa: F401

__all__ = ["Context", "Library", "PPFError"]
