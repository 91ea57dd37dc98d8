"""GPU tests of the sensitivity path: the torch.autograd.Function (forward solve + hand-written adjoint)
against the oracle's composed gradient and against central finite differences, and the
TopologyOptimizationProblem + OCOptimizer API against the oracle's restatement.

FD check (MBB half-beam 128x64x64, linear hat filter rmin=2.5, projection beta=8, SIMP p=3): 20
seeded random directions v with N(0,1) entries, central differences with h=1e-6, solves to a relative
residual of 1e-12, in FP64 and with the FP32 V-cycle. For every direction |FD - g.v| / |g.v| must be
<= 1e-6 (and |FD - g.v| / ||g||_2 as well); the printout lists both per direction.
"""

import numpy as np
import pytest

from oracle import oracle as orc
from voxelfem_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vf():
    import voxelfem_amd

    return voxelfem_amd


def mbb_pipeline(vf, ne, precision="fp64", tol=1e-12):
    from voxelfem_amd.autograd import CompliancePipeline

    tps = wl.configure(vf.pyVoxelFEM, ne, wl.MBB_BC, None, symmetry_x=True)
    return CompliancePipeline(tps, vf.optMGLevel(ne, 3), rmin=2.5, filter_type=1, beta=8.0, tol=tol,
                              max_iter=3000, precision=precision)


def oracle_mbb_gradient(ne, x, rmin=2.5, beta=8.0):
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_dirichlet(wl.mbb_mask(ne))
    s.set_forces(wl.mbb_load(ne))
    chain = orc.Chain(ne, rmin, True, beta)
    _, p = chain.forward(x)
    s.set_densities(p)
    from voxelfem_amd.mg_levels import optMGLevel

    mg = orc.MG(s, optMGLevel(ne, 3))
    f = s.build_load()
    u, _ = mg.pcg(np.zeros_like(f), f, 3000, 1e-12)
    c = 0.5 * np.sum(f * u)
    return c, chain.backprop(s.compliance_gradient(u), x)


def test_bc_masks_mbb(vf):
    ne = (16, 8, 8)
    tps = wl.configure(vf.pyVoxelFEM, ne, wl.MBB_BC, None, symmetry_x=True)
    m = tps.getDirichletMask()
    mm = m[:, 0].astype(np.uint8) | (m[:, 1].astype(np.uint8) << 1) | (m[:, 2].astype(np.uint8) << 2)
    assert np.array_equal(mm, wl.mbb_mask(ne))
    assert np.allclose(tps.buildLoadVector(), wl.mbb_load(ne))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autograd_matches_oracle(vf, precision):
    import torch

    ne = (32, 16, 16)
    x0 = np.random.default_rng(2).uniform(0.2, 0.4, int(np.prod(ne)))
    pipe = mbb_pipeline(vf, ne, precision)
    x = torch.tensor(x0, device="cuda:0", requires_grad=True)
    c = pipe(x)
    c.backward()
    cref, gref = oracle_mbb_gradient(ne, x0)
    assert abs(c.item() - cref) <= 1e-11 * abs(cref)
    g = x.grad.cpu().numpy()
    assert np.abs(g - gref).max() <= 1e-9 * np.abs(gref).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autograd_central_fd(vf, precision):
    import torch

    ne = (128, 64, 64)
    n = int(np.prod(ne))
    x0 = np.random.default_rng(2).uniform(0.2, 0.4, n)
    pipe = mbb_pipeline(vf, ne, precision, tol=1e-12)
    x = torch.tensor(x0, device="cuda:0", requires_grad=True)
    c = pipe(x)
    c.backward()
    g = x.grad.detach()
    its0 = pipe.last_stats.iterations
    rng = np.random.default_rng(20)
    h = 1e-6
    gnorm = float(torch.linalg.vector_norm(g).item())
    errs, errs_dir = [], []
    for k in range(20):
        # each design variable is perturbed by h * N(0, 1)
        vt = torch.tensor(rng.normal(size=n), device="cuda:0")
        cp = pipe(x.detach() + h * vt).item()
        cm = pipe(x.detach() - h * vt).item()
        fd = (cp - cm) / (2 * h)
        gv = float(torch.dot(g, vt).item())
        errs.append(abs(fd - gv) / gnorm)
        errs_dir.append(abs(fd - gv) / abs(gv))
        print(f"  dir {k:2d}: g.v={gv:+.6e} fd={fd:+.6e} |fd-g.v|/||g||={errs[-1]:.2e} "
              f"|fd-g.v|/|g.v|={errs_dir[-1]:.2e} its={pipe.last_stats.iterations} (base {its0})")
    print(f"FD check {precision}: ||g||={gnorm:.4e} max |fd-g.v|/||g|| {max(errs):.3e}; |fd-g.v|/|g.v| "
          f"max {max(errs_dir):.3e} median {np.median(errs_dir):.3e}")
    assert max(errs_dir) <= 1e-6
    assert max(errs) <= 1e-6


def test_problem_oc_steps_match_oracle(vf):
    pv = vf.pyVoxelFEM
    ne = (32, 16, 16)
    n = int(np.prod(ne))
    vfrac = 0.3
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, np.full(n, vfrac))
    mg = tps.multigridSolver(vf.optMGLevel(ne, 3))
    obj = pv.MultigridComplianceObjective(mg)
    obj.tol = 1e-12
    obj.cgIter = 1000
    sf = pv.SmoothingFilter(2, pv.SmoothingFilter.Type.Linear)
    top = pv.TopologyOptimizationProblem(tps, obj, [pv.TotalVolumeConstraint(vfrac)], [sf])
    top.setVars(np.full(n, vfrac))
    oc = pv.OCOptimizer(top)

    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_dirichlet(wl.cantilever_mask(ne))
    f = wl.cantilever_load(ne)
    s.set_forces(f)
    om = orc.MG(s, vf.optMGLevel(ne, 3))
    chain = orc.Chain(ne, 3.0, True, 0.0)
    ocr = orc.OC()
    xr = np.full(n, vfrac)
    u = np.zeros_like(f)

    def oracle_state(x, u0):
        _, p = chain.forward(x)
        s.set_densities(p)
        uu, _ = om.pcg(u0, f, 1000, 1e-12, 1, 2, True)
        return uu

    u = oracle_state(xr, u)
    for step in range(3):
        assert top.evaluateObjective() == pytest.approx(0.5 * np.sum(f * u), rel=1e-11)
        gref = chain.backprop(s.compliance_gradient(u), xr)
        assert np.abs(top.evaluateObjectiveGradient() - gref).max() <= 1e-9 * np.abs(gref).max()
        dc = chain.backprop(np.full(n, -1.0 / (vfrac * n)), xr)
        assert np.allclose(top.evaluateConstraintsJacobian()[0], dc, rtol=1e-12, atol=1e-18)
        oc.step()
        xr = ocr.step(xr, gref, dc, vfrac, chain)
        assert np.abs(top.getVars() - xr).max() <= 1e-9, step
        u = oracle_state(xr, u)
    assert abs(top.evaluateConstraints()[0]) <= 1e-6
