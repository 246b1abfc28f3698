"""vortrace_amd: vortrace's ray-projection path on MI355X (gfx950).

Drop-in for ``import vortrace as vt``::

    import vortrace_amd as vt
    pc = vt.ProjectionCloud(pos, rho)
    img = pc.grid_projection(extent, 512, bounds, reduction='sum')

The walk, the kNN queries and the reductions run in hand-written HIP kernels
(libvortrace_hip.so, C-ABI in include/vortrace_hip.h); this package is the
host side.  Importing it without the built library raises ImportError.
"""
import logging

from . import _lib
from . import autograd
from . import grid
from . import tuning
from .core import (PointCloud, Projection, Ray, ReductionMode, Slice, get_verbose,
                   set_verbose)
from .vortrace import ProjectionCloud

_lib.lib()  # fail loudly at import if the native library is missing

logging.getLogger("vortrace").addHandler(logging.NullHandler())

__version__ = "0.1.0"
__all__ = ["ProjectionCloud", "PointCloud", "Projection", "Ray", "ReductionMode", "Slice",
           "autograd", "grid", "tuning", "set_verbose", "get_verbose", "__version__"]
