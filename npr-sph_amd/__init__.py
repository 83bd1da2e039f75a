"""npr-sph_amd: MI355X-native SPH solver for NPR-SPH's three compute passes.

The product is libnprsph.so (HIP kernels for gfx950 behind the C ABI in
include/sph.h).  This package holds its sources (csrc/), its build script
(build.py), the ctypes host binding (sph.py) and the synthetic scene
definitions (scenes.py).

The directory name is not a Python identifier; callers load it through
importlib under the name `npr_sph_amd` (see `load_package` in
__graft_entry__.py).
"""
from . import scenes, sph  # noqa: F401
from .sph import Simulation, SphError  # noqa: F401
