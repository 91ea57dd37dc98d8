"""Stacked layer-by-layer batches against the sequential evaluator.

vf_lbl_run_zero_batched lays the partial structures of a batch side by side along x, separated by one
coarsest-level element of void, and solves them as one block-diagonal problem with per-structure PCG
scalars (LayerByLayer.hh:223-296 with the zero guess). Each structure's solve is then its own solve up to
the rounding of the PCG reductions, so objective, gradient, per-layer compliances and per-layer iteration
counts must match the sequential evaluator's (which the oracle pins in test_gpu_parity / test_gpu_tiled)
to within that rounding, over batch sizes, precisions, level counts (including a coarsest level of one
coarsening), layer increments that leave a short last layer, full multigrid, RAMP, design forces on rows that
most structures leave detached (structures without load stop before the first iteration), and constraints
on a side face (replicated into every segment).
"""
import numpy as np
import pytest

from voxelfem_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pv():
    import voxelfem_amd

    return voxelfem_amd.pyVoxelFEM


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _lbl_sim(pv, ne, seed, law=None):
    rho = wl.noise_density(ne, seed)
    full = wl.configure(pv, ne, None, rho)
    return full.getIntermediateFabricationShape(1.0, False, law if law is not None else pv.InterpolationLaw.SIMP)


def _run(pv, sim, levels, inc, prec, batched, pipes=0, graphs=True, fmg=False, tol=1e-10):
    mg = sim.multigridSolver(levels)
    ev = pv.LayerByLayerEvaluator(sim)
    ev.selectInitMethod("zero")
    ev.precision = prec
    ev.batched, ev.pipelines, ev.useGraphs = batched, pipes, graphs
    ev.run(mg, True, inc, 500, tol, fullMultigrid=fmg)
    return ev.objective(), ev.gradient(), list(ev.perLayerCompliance()), list(ev.perLayerIterations())


def _check(seq, bat, what):
    obj_s, g_s, c_s, it_s = seq
    obj_b, g_b, c_b, it_b = bat
    print(f"{what}: objective {abs(obj_b - obj_s) / abs(obj_s):.2e}, gradient {rel(g_b, g_s):.2e}, "
          f"iterations {sum(it_b)} / {sum(it_s)}", flush=True)
    assert np.isfinite(obj_b)
    assert obj_b == pytest.approx(obj_s, rel=1e-12), what
    assert rel(g_b, g_s) < 1e-11, what
    assert np.allclose(c_b, c_s, rtol=1e-11, atol=1e-300), what
    assert it_b == it_s, what


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stacked_batches_match_sequential(pv, prec):
    ne = (16, 12, 8)
    sim = _lbl_sim(pv, ne, 7)
    seq = _run(pv, sim, 2, 1, prec, False)
    for pipes, graphs in ((0, True), (5, False), (1, True)):
        _check(seq, _run(pv, sim, 2, 1, prec, True, pipes, graphs), f"{prec} batches of {pipes} graphs={graphs}")


@pytest.mark.parametrize("levels,inc", [(1, 2), (3, 3)])
def test_stacked_levels_and_short_last_layer(pv, levels, inc):
    ne = (16, 8, 8) if levels == 3 else (8, 6, 4)
    sim = _lbl_sim(pv, ne, 11)
    seq = _run(pv, sim, levels, inc, "fp64", False)
    _check(seq, _run(pv, sim, levels, inc, "fp64", True, 2), f"L={levels} increment {inc}")


def test_stacked_full_multigrid_and_ramp(pv):
    ne = (16, 8, 8)
    sim = _lbl_sim(pv, ne, 5, pv.InterpolationLaw.RAMP)
    seq = _run(pv, sim, 2, 1, "fp32", False, fmg=True)
    _check(seq, _run(pv, sim, 2, 1, "fp32", True, 3, True, fmg=True), "RAMP, full multigrid")


def test_stacked_forces_on_detached_rows_and_side_constraints(pv):
    """A load on the design's top face loads only the full structure (the others leave that row detached and
    stop before the first iteration with zero compliance); the x = 0 face is clamped in every segment."""
    ne = (16, 8, 8)
    sim = _lbl_sim(pv, ne, 3)
    sim.gravity = [0.0, 0.0, 0.0]
    X = wl._node_positions(ne)
    f = np.zeros((len(X), 3))
    top = X[:, 1] == ne[1]
    f[top, 1] = -1.0 / top.sum()
    sim.setForces(f)
    sim.addDirichletCondition([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, float(ne[1]), float(ne[2])], "xyz")
    seq = _run(pv, sim, 2, 1, "fp64", False)
    assert seq[2][0] > 0 and all(c == 0 for c in seq[2][1:])
    _check(seq, _run(pv, sim, 2, 1, "fp64", True, 0, True), "top-face load, clamped side")


def test_stacked_batches_reused_across_evaluations(pv):
    """The solver keeps its stacked batches between evaluations: after new densities (same batches, new
    moduli and hierarchy) and after a new constraint (batches rebuilt) the results still match."""
    ne = (16, 8, 8)
    sim = _lbl_sim(pv, ne, 21)
    mg = sim.multigridSolver(2)
    ev = pv.LayerByLayerEvaluator(sim)
    ev.selectInitMethod("zero")

    def both(what):
        out = []
        for batched in (False, True):
            ev.batched, ev.pipelines = batched, 3
            ev.run(mg, True, 1, 500, 1e-10)
            out.append((ev.objective(), ev.gradient(), list(ev.perLayerCompliance()), list(ev.perLayerIterations())))
        _check(out[0], out[1], what)
        return out[1][0]

    o1 = both("first evaluation")
    sim.setDensities(wl.noise_density(ne, 22))
    o2 = both("new densities")
    sim.addDirichletCondition([0.0, 0.0, 0.0], [float(ne[0]), 0.0, 0.0], [float(ne[0]), float(ne[1]), float(ne[2])],
                              "x")
    o3 = both("new constraint")
    assert o1 != o2 and o2 != o3
