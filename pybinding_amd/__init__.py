"""pybinding_amd -- an MI355X-native kernel polynomial method engine.

A drop-in compute backend for the reference's `pb.kpm(model, ...)` (pybinding/chebyshev.py):
the host C++ core (`_kpmx`) calls hand-written HIP kernels (`libkpmx.so`, C-ABI in
include/kpmx.h) for every Chebyshev moment.
"""
from .chebyshev import *  # noqa: F401,F403
from .chebyshev import KPM, kpm, kpm_cuda, kpm_hip, jackson_kernel, lorentz_kernel, dirichlet_kernel  # noqa: F401
from .results import Series, StructureMap, Deferred  # noqa: F401
from .io import save, load  # noqa: F401
from . import generators  # noqa: F401
from .model import Model  # noqa: F401
from .lattice import (Lattice, Polygon, rectangle, regular_polygon, primitive, translational_symmetry,  # noqa: F401
                      onsite_energy_modifier, hopping_energy_modifier, site_position_modifier,
                      constant_potential, force_double_precision)

__version__ = "0.1.0"
