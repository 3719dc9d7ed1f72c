"""framework_amd — MI355X-native Dual Marching Cubes.

The product is the HIP library ``framework_amd/lib/libmdmc.so`` behind the C ABI
in ``include/mdmc.h`` (a drop-in for DMC_CUDA's dmc::run_dmc path).  This
package holds its sources (``csrc/``), the build step (``build.py``) and thin
Python bindings (``mdmc.py``).
"""
from .mdmc import (Context, MDMCError, cleanup, declared_functions, lib, run_dmc, run_host, setup,
                   triangles_from_quads, write_obj, run_out_of_core, out_of_core_slab_planes, run_device)

__all__ = ["Context", "MDMCError", "cleanup", "declared_functions", "lib", "run_dmc", "run_host", "setup",
           "triangles_from_quads", "write_obj", "run_out_of_core",
           "out_of_core_slab_planes", "run_device"]
