"""Parity of the HIP path with the oracle on the benchmark workloads at their full sizes.

  C1  256x128x128 cantilever, smoothed density: FP64 PCG with the FP64 V-cycle (u to 1e-10 relative,
      compliance to 1e-12) and with the FP32 V-cycle (compliance to 1e-9), both to 1e-8 residual.
  C3  160x80x80 layer-by-layer, 40 layers of 2 along the build axis, clamped base, self-weight: the
      batched and the sequential evaluators against the oracle's 40 solves (objective to 1e-9,
      gradient to 1e-8 of its largest entry).
  C4  384x192x192 cantilever (43,022,595 DoF), uniform density 0.3: one solve in FP64 and with the FP32
      V-cycle against the oracle's full solve, and the first three optimality-criterion iterations of the
      design loop (filter radius 3, volume fraction 0.3) against the oracle's OC steps.

The oracle runs multi-threaded on the host; these tests take tens of seconds each.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from voxelfem_amd import workloads as wl
from voxelfem_amd.mg_levels import optMGLevel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pv():
    import voxelfem_amd

    return voxelfem_amd.pyVoxelFEM


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def oracle_cantilever(ne, rho):
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    s.set_dirichlet(wl.cantilever_mask(ne))
    s.set_forces(wl.cantilever_load(ne))
    return s


def _solve_both(pv, ne, rho):
    levels = optMGLevel(ne, 3)
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, rho)
    mg = tps.multigridSolver(levels)
    f = tps.buildLoadVector()
    u0 = np.zeros_like(f)
    u64 = mg.preconditionedConjugateGradient(u0, f, 1000, 1e-8)
    u32 = mg.preconditionedConjugateGradient(u0, f, 1000, 1e-8, precision="fp32")
    r = f - tps.applyK(u32)
    r[tps.getDirichletMask()] = 0
    res32 = np.linalg.norm(r) / np.linalg.norm(f)
    print(f"{ne}: GPU solves done; oracle solve", flush=True)
    s = oracle_cantilever(ne, rho)
    ref, its = orc.MG(s, levels).pcg(u0, f, 1000, 1e-8)
    c64, c32, cref = (0.5 * float(np.sum(f * v)) for v in (u64, u32, ref))
    print(f"{ne}: oracle {its} iterations; FP64 |du|/|u| {rel(u64, ref):.2e}, compliance {abs(c64 - cref) / cref:.2e}; "
          f"FP32 V-cycle compliance {abs(c32 - cref) / cref:.2e}, residual {res32:.2e}")
    assert rel(u64, ref) <= 1e-10
    assert abs(c64 - cref) <= 1e-12 * abs(cref)
    assert abs(c32 - cref) <= 1e-9 * abs(cref)
    assert res32 <= 1e-8 * 1.0001


@pytest.mark.timeout(600)
def test_c1_fp64_and_fp32_vcycle_match_oracle(pv):
    ne = wl.GRIDS["cantilever_256x128x128"]
    _solve_both(pv, ne, wl.smoothed_density(ne, 1))


@pytest.mark.timeout(600)
def test_c4_solve_matches_oracle(pv):
    ne = wl.GRIDS["cantilever_384x192x192"]
    _solve_both(pv, ne, np.full(int(np.prod(ne)), 0.3))


@pytest.mark.timeout(900)
def test_c3_layer_by_layer_matches_oracle(pv):
    ne = wl.GRIDS["lbl_160x80x80"]
    rho = wl.smoothed_density(ne, 3)
    levels = optMGLevel(ne, 3)
    full = wl.configure(pv, ne, None, rho)
    sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = sim.multigridSolver(levels)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    X = wl._node_positions(ne)
    s.set_dirichlet(np.where(X[:, 1] == 0, 7, 0).astype(np.uint8))
    s.set_gravity([0.0, -1.0, 0.0])
    print("C3: oracle layer-by-layer run (40 solves)", flush=True)
    obj, grad, comp, its = orc.MG(s, levels).lbl_run_zero(2, 1000, 1e-8)
    print("C3: oracle done", flush=True)
    assert len(comp) == 40
    for batched in (True, False):
        ev = pv.LayerByLayerEvaluator(sim)
        ev.selectInitMethod("zero")
        ev.batched = batched
        ev.run(mg, True, 2, 1000, 1e-8)
        print(f"C3 batched={batched}: objective {abs(ev.objective() - obj) / obj:.2e}, gradient "
              f"{rel(ev.gradient(), grad):.2e}, {sum(ev.perLayerIterations())} PCG iterations (oracle {sum(its)})",
              flush=True)
        assert ev.objective() == pytest.approx(obj, rel=1e-9)
        assert rel(ev.gradient(), grad) <= 1e-8
        assert np.allclose(ev.perLayerCompliance(), comp, rtol=1e-9, atol=0)


@pytest.mark.timeout(900)
def test_c4_oc_iterations_match_oracle(pv):
    """The first three iterations of the 384x192x192 design loop: objective (multigrid compliance, FMG
    V-cycle with 2+2 sweeps, warm start, tol 1e-8), filtered gradient and the OC update (m = 0.2, p = 0.5,
    ctol = 1e-6) against the oracle's restatement of the same loop."""
    ne = wl.GRIDS["cantilever_384x192x192"]
    n = int(np.prod(ne))
    vfrac = 0.3
    levels = optMGLevel(ne, 3)
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, np.full(n, vfrac))
    mg = tps.multigridSolver(levels)
    obj = pv.MultigridComplianceObjective(mg)
    obj.tol = 1e-8
    obj.cgIter = 1000
    top = pv.TopologyOptimizationProblem(tps, obj, [pv.TotalVolumeConstraint(vfrac)],
                                         [pv.SmoothingFilter(3, pv.SmoothingFilter.Type.Linear)])
    top.setVars(np.full(n, vfrac))
    oc = pv.OCOptimizer(top)

    s = oracle_cantilever(ne, np.full(n, vfrac))
    f = wl.cantilever_load(ne)
    om = orc.MG(s, levels)
    chain = orc.Chain(ne, 4.0, True, 0.0)  # reference radius 3 = hat of support 4
    ocr = orc.OC()
    xr = np.full(n, vfrac)

    def oracle_state(x, u0):
        _, p = chain.forward(x)
        s.set_densities(p)
        # the objective solves once at construction and again after setVars, both warm-started
        uu, its = om.pcg(u0, f, 1000, 1e-8, 1, 2, True)
        print(f"C4 OC: oracle solve, {its} iterations", flush=True)
        return uu

    u = oracle_state(xr, np.zeros_like(f))
    u = oracle_state(xr, u)
    for step in range(3):
        c, cref = top.evaluateObjective(), 0.5 * float(np.sum(f * u))
        g = top.evaluateObjectiveGradient()
        gref = chain.backprop(s.compliance_gradient(u), xr)
        dc = chain.backprop(np.full(n, -1.0 / (vfrac * n)), xr)
        oc.step()
        xr = ocr.step(xr, gref, dc, vfrac, chain)
        dx = np.abs(top.getVars() - xr).max()
        print(f"C4 OC step {step}: compliance {abs(c - cref) / cref:.2e}, gradient {rel(g, gref):.2e}, "
              f"max |x - x_oracle| {dx:.2e}", flush=True)
        assert c == pytest.approx(cref, rel=1e-10)
        assert rel(g, gref) <= 1e-8
        assert dx <= 1e-8, step
        u = oracle_state(xr, u)
    assert abs(top.evaluateConstraints()[0]) <= 1e-6
