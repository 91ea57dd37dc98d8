"""CPU-side checks of the native boundary: the C ABI library loads and exports every entry point
that include/voxelfem.h declares, and the pybind module exposes the reference's API surface.
No device work is done here (there is no GPU in the CPU test environment)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "voxelfem.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(vf_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "vf_mg_pcg" in syms and "vf_sim_apply_K" in syms and len(syms) > 40


def test_library_exports_every_declared_symbol():
    import voxelfem_amd

    lib = ctypes.CDLL(voxelfem_amd.LIBRARY_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    lib.vf_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.vf_version()


def test_pybind_surface():
    from voxelfem_amd import pyVoxelFEM as pv

    for name in ["TensorProductSimulator", "TopologyOptimizationProblem", "ComplianceObjective",
                 "MultigridComplianceObjective", "LayerByLayerEvaluator", "OCOptimizer", "getClassName",
                 "FilterChain", "SmoothingFilter", "ProjectionFilter", "PythonFilter", "TotalVolumeConstraint",
                 "InterpolationLaw", "NumberType", "UpsampleFilter", "VertexToCellFilter", "LangelaarFilter"]:
        assert hasattr(pv, name), name
    tps_cls = pv.detail.TensorProductSimulator1_1_1
    for meth in ["numNodes", "numElements", "getDensities", "setDensities", "setUniformDensities", "applyK",
                 "multigridSolver", "complianceGradient", "buildLoadVector", "applyDisplacementsAndLoadsFromFile",
                 "readMaterial", "addDirichletCondition", "applySymmetryConditions", "getIntermediateFabricationShape",
                 "setFabricationMaskHeightByLayer", "elementEnergyDensity", "solve"]:
        assert hasattr(tps_cls, meth), meth
    mg_cls = pv.detail.MultigridSolver1_1_1
    for meth in ["preconditionedConjugateGradient", "solve", "applyK", "computeResidual",
                 "zeroOutDirichletComponents", "updateStiffnessMatrices", "setSymmetricGaussSeidel",
                 "setFabricationMaskHeightByLayer", "getSimulator"]:
        assert hasattr(mg_cls, meth), meth
    assert pv.getClassName(None, "TensorProductSimulator") == "pyVoxelFEM.detail.TensorProductSimulator1_1_1"
    with pytest.raises(RuntimeError, match="No template instantiation"):
        pv.TensorProductSimulator([2, 2, 2], [[0, 0, 0], [1, 1, 1]], [2, 2, 2])


def test_level_heuristic_matches_survey_counts():
    from voxelfem_amd.mg_levels import optMGLevel, possibleMaxMGLevel

    assert [optMGLevel(d, 3) for d in [(64, 32, 32), (256, 128, 128), (128, 64, 64), (160, 80, 80),
                                        (384, 192, 192)]] == [2, 4, 3, 3, 5]
    assert possibleMaxMGLevel((160, 80, 80)) == 4
    assert optMGLevel((8, 4, 4), 3) == 1 and optMGLevel((1, 4, 4), 3) == 0


def test_grid_changing_filter_dimensions():
    """UpsampleFilter / VertexToCellFilter / LangelaarFilter dimension rules and their errors
    (TopologyOptimizationFilter.hh:441-456, 587-595), checked without touching a device."""
    from voxelfem_amd import pyVoxelFEM as pv

    u = pv.UpsampleFilter(3)
    u.setInputDimensions([4, 5])
    assert (u.inputDimensions, u.outputDimensions) == ([4, 5], [10, 13])
    u.setOutputDimensions([7, 4, 13])
    assert (u.inputDimensions, u.outputDimensions) == ([3, 2, 5], [7, 4, 13])
    with pytest.raises(RuntimeError, match="not divisible by factor"):
        u.setOutputDimensions([11, 13])
    with pytest.raises(RuntimeError, match="2\\^d grid or larger"):
        u.setInputDimensions([1, 5])
    v = pv.VertexToCellFilter()
    v.setOutputDimensions([6, 7])
    assert (v.inputDimensions, v.outputDimensions) == ([7, 8], [6, 7])
    with pytest.raises(RuntimeError, match="2\\^d or larger"):
        v.setInputDimensions([1, 3, 3])
    lg = pv.LangelaarFilter()
    lg.setOutputDimensions([5, 4])
    assert lg.inputDimensions == [5, 4]
    with pytest.raises(RuntimeError, match="planar grids"):
        lg.setOutputDimensions([4, 4, 4])
    with pytest.raises(RuntimeError, match="not set"):
        pv.VertexToCellFilter().apply(np.zeros(4))
