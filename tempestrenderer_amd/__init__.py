"""MI355X (gfx950) path tracer behind the reference's ray-trace scene API.

The compute path is libtempest_rt.so (HIP kernels + C ABI, built in-tree);
this package holds the host-side mirror of the scene API and the synthetic
workload generators.
"""
from ._abi import TempestError, load  # noqa: F401
from .scene import Material, RayTracerScene, Submesh, make_light  # noqa: F401
