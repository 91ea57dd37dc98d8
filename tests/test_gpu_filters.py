"""Parity of the grid-changing and build-direction filters (UpsampleFilter, VertexToCellFilter, LangelaarFilter;
TopologyOptimizationFilter.hh:419-712) on the device with their numpy restatement (oracle/filters.py, pinned in
tests/test_filters_oracle.py).

Tolerances: upsample and vertex-to-cell to 1e-13 relative (the kernels may contract a multiply-add); Langelaar
filtered densities to 1e-12 and its backprop to 1e-10 relative (powers by repeated squaring and device pow against
the host's pow, carried through the
layer recursion).
"""
import ctypes

import numpy as np
import pytest

from oracle import filters as F
from voxelfem_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pv():
    import voxelfem_amd

    return voxelfem_amd.pyVoxelFEM


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def chain_backprop(pv, filters, out_dims, x, g):
    fc = pv.FilterChain(filters, list(out_dims))
    fc.setDesignVars(x.ravel())
    return fc, fc.backprop(g.ravel())


@pytest.mark.parametrize("shape,factor", [((5, 4), 2), ((3, 7), 3), ((4, 3, 5), 2), ((2, 3, 2), 4), ((6, 5), 1),
                                          ((30, 70), 3), ((6, 11, 40), 2)])
def test_upsample_matches_oracle(pv, shape, factor):
    rng = np.random.default_rng(shape[0] * 31 + factor)
    x = rng.random(shape)
    out_dims = F.upsample_dims(shape, factor)
    f = pv.UpsampleFilter(factor)
    f.setInputDimensions(list(shape))
    assert f.outputDimensions == list(out_dims)
    y = f.apply(x.ravel())
    assert rel(y, F.upsample(x, factor).ravel()) < 1e-13
    g = rng.standard_normal(out_dims)
    fc, gin = chain_backprop(pv, [pv.UpsampleFilter(factor)], out_dims, x, g)
    assert fc.gridDims() == list(shape) and fc.physicalGridDims() == list(out_dims)
    assert fc.numVars() == x.size
    assert rel(fc.physicalVars(), F.upsample(x, factor).ravel()) < 1e-13
    assert rel(gin, F.upsample_backprop(g, shape, factor).ravel()) < 1e-13


@pytest.mark.parametrize("shape", [(5, 4), (2, 9), (4, 3, 5), (2, 2, 2), (41, 130), (7, 9, 70)])
def test_vertex_to_cell_matches_oracle(pv, shape):
    rng = np.random.default_rng(shape[0] * 7 + len(shape))
    x = rng.random(shape)
    out_dims = tuple(n - 1 for n in shape)
    f = pv.VertexToCellFilter()
    f.setOutputDimensions(list(out_dims))
    assert f.inputDimensions == list(shape)
    assert rel(f.apply(x.ravel()), F.vertex_to_cell(x).ravel()) < 1e-13
    g = rng.standard_normal(out_dims)
    _, gin = chain_backprop(pv, [pv.VertexToCellFilter()], out_dims, x, g)
    assert rel(gin, F.vertex_to_cell_backprop(g, shape).ravel()) < 1e-13


def test_c_abi_host_buffers(pv):
    """vf_upsample_filter / vf_vertex_to_cell with host buffers, called directly through the C ABI."""
    from voxelfem_amd import capi

    L = capi.lib()
    shape, factor = (4, 6, 3), 3
    x = np.random.default_rng(5).random(shape)
    dims = (ctypes.c_int64 * 3)(*shape)
    up = np.empty(F.upsample_dims(shape, factor))
    capi.check(L.vf_upsample_filter(3, dims, factor, x.ctypes.data, up.ctypes.data, capi.VF_HOST, None))
    assert rel(up, F.upsample(x, factor)) < 1e-13
    back = np.empty(shape)
    capi.check(L.vf_upsample_filter_backprop(3, dims, factor, up.ctypes.data, back.ctypes.data, capi.VF_HOST, None))
    assert rel(back, F.upsample_backprop(up, shape, factor)) < 1e-13
    cells = np.empty(tuple(n - 1 for n in shape))
    capi.check(L.vf_vertex_to_cell(3, dims, x.ctypes.data, cells.ctypes.data, capi.VF_HOST, None))
    assert rel(cells, F.vertex_to_cell(x)) < 1e-13
    verts = np.empty(shape)
    capi.check(L.vf_vertex_to_cell_backprop(3, dims, cells.ctypes.data, verts.ctypes.data, capi.VF_HOST, None))
    assert rel(verts, F.vertex_to_cell_backprop(cells, shape)) < 1e-13
    bad = (ctypes.c_int64 * 3)(1, 4, 4)
    assert L.vf_vertex_to_cell(3, bad, x.ctypes.data, cells.ctypes.data, capi.VF_HOST, None) != 0
    assert b"too small" in L.vf_last_error()


@pytest.mark.parametrize("shape", [(9, 7), (1, 5), (40, 24), (2048, 3), (2100, 4)])
def test_langelaar_matches_oracle(pv, shape):
    """Grids up to 2048 columns run the single-workgroup recurrence, wider ones one launch per layer."""
    rng = np.random.default_rng(shape[0] + 100 * shape[1])
    x = 0.05 + 0.9 * rng.random(shape)
    filtered, smax = F.langelaar(x)
    f = pv.LangelaarFilter()
    f.setInputDimensions(list(shape))
    assert rel(f.apply(x.ravel()), filtered.ravel()) < 1e-12
    g = rng.standard_normal(shape)
    fc, gin = chain_backprop(pv, [pv.LangelaarFilter()], shape, x, g)
    assert rel(fc.physicalVars(), filtered.ravel()) < 1e-12
    assert rel(gin, F.langelaar_backprop(g, x, filtered, smax).ravel()) < 1e-10


def test_langelaar_overhang(pv):
    """The device filter removes an unsupported ledge as the oracle does."""
    x = np.zeros((8, 7))
    x[3, :] = 1.0
    x[1:7, 4] = 1.0
    f = pv.LangelaarFilter()
    f.setOutputDimensions([8, 7])
    y = f.apply(x.ravel()).reshape(8, 7)
    assert rel(y, F.langelaar(x)[0]) < 1e-12
    assert y[1, 4] < 0.05 and y[2, 4] > 0.95


def test_mixed_chain_matches_oracle(pv):
    """Upsample -> vertex-to-cell -> projection on a planar grid: design grid 5 x 4 vertices, physical 8 x 6
    cells; values and chained adjoint."""
    out_dims = (8, 6)
    fc = pv.FilterChain([pv.UpsampleFilter(2), pv.VertexToCellFilter(), pv.ProjectionFilter(4.0)], list(out_dims))
    assert fc.gridDims() == [5, 4] and fc.physicalGridDims() == [8, 6] and fc.numVars() == 20
    rng = np.random.default_rng(11)
    x = rng.random((5, 4))
    fc.setDesignVars(x.ravel())
    up = F.upsample(x, 2)
    cells = F.vertex_to_cell(up)
    beta = 4.0
    proj = (np.tanh(beta * 0.5) + np.tanh(beta * (cells - 0.5))) / (2 * np.tanh(beta * 0.5))
    assert rel(fc.physicalVars(), proj.ravel()) < 1e-13
    g = rng.standard_normal(out_dims)
    dproj = beta * (1 - np.tanh(beta * (cells - 0.5)) ** 2) / (2 * np.tanh(beta * 0.5))
    expect = F.upsample_backprop(F.vertex_to_cell_backprop(g * dproj, up.shape), (5, 4), 2)
    assert rel(fc.backprop(g.ravel()), expect.ravel()) < 1e-12


def test_problem_with_grid_changing_filters(pv):
    """A planar compliance problem whose design grid is coarser than its elements: the objective gradient is the
    chain's adjoint of the element sensitivities, and an OC step keeps the design-grid size."""
    ne = (16, 8)
    rho = wl.noise_density(ne, 3)
    tps = wl.configure2d(pv, ne, wl.CANTILEVER2D_BC, rho)
    obj = pv.MultigridComplianceObjective(tps.multigridSolver(3))
    obj.tol, obj.cgIter = 1e-10, 200
    filters = [pv.UpsampleFilter(2), pv.VertexToCellFilter(), pv.LangelaarFilter()]
    problem = pv.TopologyOptimizationProblem(tps, obj, [pv.TotalVolumeConstraint(0.4)], filters)
    assert problem.numVars() == 9 * 5
    x = 0.3 + 0.4 * np.random.default_rng(2).random(45)
    problem.setVars(x)
    dJ_phys = obj.gradient()
    fc = problem.filterChain
    up = F.upsample(x.reshape(9, 5), 2)
    cells = F.vertex_to_cell(up)
    filtered, smax = F.langelaar(cells)
    assert rel(fc.physicalVars(), filtered.ravel()) < 1e-12
    gl = F.langelaar_backprop(dJ_phys.reshape(ne), cells, filtered, smax)
    expect = F.upsample_backprop(F.vertex_to_cell_backprop(gl, up.shape), (9, 5), 2)
    assert rel(problem.evaluateObjectiveGradient(), expect.ravel()) < 1e-10
    jac = problem.evaluateConstraintsJacobian()
    assert jac.shape == (1, 45)
    opt = pv.OCOptimizer(problem)
    opt.step()
    assert problem.getVars().shape == (45,)
    assert abs(problem.evaluateConstraints()[0]) < 1e-5
