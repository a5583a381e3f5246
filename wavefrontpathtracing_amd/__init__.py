"""MI355X wavefront path tracer: HIP kernels behind a C-ABI (include/wf.h).

The package holds the native library (libwf.so, built by ``build.py``), its
ctypes binding (``_lib``), the Python mirror of the reference renderer API
(``renderer``) and the procedural workloads (``scenes``).
"""
from ._lib import HIT, RAY, VERTEX, WfError, lib  # noqa: F401
from .renderer import Material, Renderer, bvh_build, camera_create, make_bvh_instance, tlas_build  # noqa: F401
