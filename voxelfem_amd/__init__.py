"""MI355X-native VoxelFEM solve-and-sensitivity path.

``voxelfem_amd.pyVoxelFEM`` is the drop-in module with the reference's Python API
(TensorProductSimulator, MultigridSolver, TopologyOptimizationProblem, filters, OC, layer-by-layer),
backed by hand-written HIP kernels for gfx950 in ``libvoxelfem.so`` (C ABI: include/voxelfem.h).

torch, when installed, is imported before the native library: torch ships its own HIP runtime and
loading it after libvoxelfem.so would put two HIP runtimes in one process. Importing it first makes
libvoxelfem.so bind to the runtime torch already loaded.
"""
import os

try:  # noqa: SIM105
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the pybind API
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libvoxelfem.so")

try:
    from . import pyVoxelFEM  # noqa: E402
except ImportError as e:  # the product has no CPU fallback: fail loudly
    raise ImportError(
        "voxelfem_amd native extension not built (run `python -m voxelfem_amd.build`): " + str(e)) from e

from .mg_levels import optMGLevel, possibleMaxMGLevel  # noqa: E402,F401

__all__ = ["pyVoxelFEM", "optMGLevel", "possibleMaxMGLevel", "LIBRARY_PATH"]
