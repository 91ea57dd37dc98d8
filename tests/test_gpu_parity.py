"""Parity of the HIP path (pyVoxelFEM -> C ABI -> gfx950 kernels) with the CPU oracle.

Tolerances: FP64 kernels agree with the oracle to rounding (relative 1e-12 on operator outputs);
the FP64 MG-PCG solution agrees to 1e-10 relative in u and 1e-12 in compliance (the reference
parity bar); the FP32-preconditioned flexible PCG reaches the same 1e-8 residual and its compliance
is within 1e-9 relative of the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from voxelfem_amd import workloads as wl

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)


@pytest.fixture(scope="module")
def pv():
    import voxelfem_amd

    return voxelfem_amd.pyVoxelFEM


def make_pair(pv, ne, rho=None, bc=True, levels=None, seed=0):
    """A pyVoxelFEM simulator and the oracle simulator with identical inputs."""
    rho = wl.noise_density(ne, seed) if rho is None else rho
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC if bc else None, rho)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    if bc:
        s.set_dirichlet(wl.cantilever_mask(ne))
        s.set_forces(wl.cantilever_load(ne))
    return tps, s


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def test_K0_and_young(pv):
    ne = (6, 4, 4)
    tps, s = make_pair(pv, ne)
    assert rel(tps.fullDensityElementStiffnessMatrix(), s.K0()) < 1e-14
    assert rel(tps.getYoungModulusScaleFactor(), s.E()) < 1e-14


def test_bc_loader_matches_definition(pv):
    ne = (8, 4, 6)
    tps, s = make_pair(pv, ne)
    m = tps.getDirichletMask()
    assert np.array_equal(m.all(1), wl.cantilever_mask(ne) == 7)
    assert np.allclose(tps.buildLoadVector(), wl.cantilever_load(ne), atol=0)


@pytest.mark.parametrize("ne", [(4, 4, 4), (6, 10, 6), (32, 16, 16), (64, 32, 32)])
def test_fine_applyK(pv, ne):
    tps, s = make_pair(pv, ne)
    u = RNG.normal(size=(tps.numNodes(), 3))
    assert rel(tps.applyK(u), s.apply_K(u)) < 1e-12


def test_fine_applyK_masked_layers(pv):
    ne = (8, 8, 8)
    tps, s = make_pair(pv, ne)
    tps.setFabricationMaskHeightByLayer(5)
    mg = orc.MG(s, 1)
    mg.set_mask_layer(5)
    u = RNG.normal(size=(tps.numNodes(), 3))
    y, ref = tps.applyK(u), s.apply_K(u)
    assert rel(y, ref) < 1e-12


@pytest.mark.parametrize("levels,variant", [(1, 0), (2, 0), (3, 0), (2, 1), (3, 1), (2, 5), (3, 5)])
def test_level_operators(pv, levels, variant):
    """Every level's K u, residual, one smoothing sweep, transfers and the coarse solve against the
    oracle. variant 0 holds level 1 as assembled block rows; variant 1 applies and smooths level 1 on the
    fly from the fine moduli on the vector units (the reflection form of the coarsened element matrices),
    variant 5 on the matrix cores (FP64 MFMA)."""
    from voxelfem_amd import capi

    ne = (16, 8, 16)
    tps, s = make_pair(pv, ne)
    capi.set_kernel_variant(variant)
    try:
        mg, om = tps.multigridSolver(levels), orc.MG(s, levels)
        mg.updateStiffnessMatrices()
    finally:
        capi.set_kernel_variant(0)
    om.update_stiffness()
    for l in range(levels + 1):
        nn = om.num_nodes(l)
        u = RNG.normal(size=(nn, 3))
        b = RNG.normal(size=(nn, 3))
        assert rel(mg.applyK(l, u), om.apply_K(l, u)) < 1e-12, ("applyK", l)
        assert rel(mg.computeResidual(l, u, b), om.residual(l, u, b)) < 1e-12, ("residual", l)
        if l < levels:
            for fwd in (True, False):
                # GPU smoothing: one sweep through the C ABI on host data
                out_gpu = _smooth(mg, l, u, b, fwd)
                out_ref = om.smooth(l, u, b, fwd)
                assert rel(out_gpu, out_ref) < 1e-11, ("smooth", l, fwd)
            nc = om.num_nodes(l + 1)
            xc = RNG.normal(size=(nc, 3))
            assert rel(_restrict(mg, l, u), om.restrict(l, u)) < 1e-13
            assert rel(_interp(mg, l, xc, b, True), om.interpolate(l, xc, b, True)) < 1e-13
    bc_ = RNG.normal(size=(om.num_nodes(levels), 3))
    assert rel(_coarse(mg, bc_), om.coarse_solve(bc_)) < 1e-9


@pytest.mark.parametrize("ne,levels", [((32, 16, 16), 2), ((24, 16, 16), 1)])
def test_coarse_solve_many_band_blocks(pv, ne, levels):
    """Coarsest direct solve whose reduced operator spans many 64-row band blocks (10 blocks with 2
    off-diagonal block bands; 47 blocks with 5), so every stage of the blocked band factorisation and
    of the strip-parallel inverse runs, against the oracle's dense Cholesky solve."""
    tps, s = make_pair(pv, ne)
    mg, om = tps.multigridSolver(levels), orc.MG(s, levels)
    mg.updateStiffnessMatrices()
    om.update_stiffness()
    for _ in range(2):
        bc_ = RNG.normal(size=(om.num_nodes(levels), 3))
        assert rel(_coarse(mg, bc_), om.coarse_solve(bc_)) < 1e-9


def test_coarse_factor_graph_replays(pv):
    """With useGraphs the coarsest factorisation is replayed from a captured HIP graph. Refactorise for
    several density fields, with eager solves and host synchronisations between the replays; every
    coarse solve must equal the eagerly launched factorisation's bit for bit and the oracle's dense
    Cholesky solve to 1e-9."""
    ne, levels = (32, 16, 16), 2
    tps, _ = make_pair(pv, ne)
    ref_tps, _ = make_pair(pv, ne)
    mg, ref = tps.multigridSolver(levels), ref_tps.multigridSolver(levels)
    mg.useGraphs, ref.useGraphs = True, False
    for seed in (3, 4, 5):
        rho = wl.noise_density(ne, seed)
        tps.setDensities(rho)
        ref_tps.setDensities(rho)
        mg.updateStiffnessMatrices()
        ref.updateStiffnessMatrices()
        s = orc.Sim(ne)
        s.set_etensor(1.0, 0.3)
        s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
        s.set_densities(rho)
        s.set_dirichlet(wl.cantilever_mask(ne))
        om = orc.MG(s, levels)
        om.update_stiffness()
        for _ in range(2):
            bc_ = RNG.normal(size=(om.num_nodes(levels), 3))
            got = _coarse(mg, bc_)
            assert np.array_equal(got, _coarse(ref, bc_)), seed
            assert rel(got, om.coarse_solve(bc_)) < 1e-9, seed


def _smooth(mg, l, u, b, fwd):
    return mg._smooth(l, u, b, fwd)


def _restrict(mg, l, u):
    return mg._restrict(l, u)


def _interp(mg, l, xc, xf, acc):
    return mg._interpolate(l, xc, xf, acc)


def _coarse(mg, b):
    return mg._coarseSolve(b)


@pytest.mark.parametrize("ne,levels,fmg,nu", [((32, 16, 16), 2, False, 1), ((32, 16, 16), 2, True, 2),
                                               ((64, 32, 32), 2, False, 1), ((32, 32, 16), 3, False, 1)])
def test_pcg_matches_oracle(pv, ne, levels, fmg, nu):
    tps, s = make_pair(pv, ne)
    mg, om = tps.multigridSolver(levels), orc.MG(s, levels)
    f = tps.buildLoadVector()
    u0 = np.zeros_like(f)
    u = mg.preconditionedConjugateGradient(u0, f, 500, 1e-8, mgIterations=1, mgSmoothingIterations=nu,
                                           fullMultigrid=fmg)
    ref, its = om.pcg(u0, f, 500, 1e-8, 1, nu, fmg)
    assert rel(u, ref) < 1e-10
    c, cref = 0.5 * np.sum(f * u), 0.5 * np.sum(f * ref)
    assert abs(c - cref) <= 1e-12 * abs(cref)


def test_pcg_fp32_preconditioner(pv):
    ne = (64, 32, 32)
    tps, s = make_pair(pv, ne)
    mg, om = tps.multigridSolver(2), orc.MG(s, 2)
    f = tps.buildLoadVector()
    u0 = np.zeros_like(f)
    its = []
    u = mg.preconditionedConjugateGradient(u0, f, 500, 1e-8, it_callback=lambda i, x, r: its.append(np.linalg.norm(r)),
                                           precision="fp32")
    ref, n64 = om.pcg(u0, f, 500, 1e-8, 1, 1, False)
    r = f - tps.applyK(u)
    r[tps.getDirichletMask()] = 0
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(f) * 1.0001
    c, cref = 0.5 * np.sum(f * u), 0.5 * np.sum(f * ref)
    assert abs(c - cref) <= 1e-9 * abs(cref)
    assert len(its) <= 2 * n64 + 5


def test_pcg_nan_raises(pv):
    tps, s = make_pair(pv, (8, 4, 4))
    mg = tps.multigridSolver(1)
    f = tps.buildLoadVector()
    f[-3, 0] = np.nan
    with pytest.raises(Exception, match="NaN encountered"):
        mg.preconditionedConjugateGradient(np.zeros_like(f), f, 10, 1e-8)


def test_compliance_gradient_and_energy(pv):
    ne = (16, 8, 8)
    tps, s = make_pair(pv, ne)
    u = RNG.normal(size=(tps.numNodes(), 3))
    assert rel(tps.complianceGradient(u), s.compliance_gradient(u)) < 1e-12
    assert rel(tps.elementEnergyDensity(u), s.energy_density(u)) < 1e-12
    tps.gravity = [0.0, -1.0, 0.0]
    s.set_gravity([0.0, -1.0, 0.0])
    assert rel(tps.buildLoadVector(), s.build_load()) < 1e-14
    assert rel(tps.complianceGradient(u), s.compliance_gradient(u)) < 1e-12


@pytest.mark.parametrize("radius,linear", [(1, True), (2, True), (3, True), (2, False)])
def test_smoothing_filter(pv, radius, linear):
    dims = (12, 7, 9)
    x = RNG.uniform(size=int(np.prod(dims)))
    sf = pv.SmoothingFilter(radius, pv.SmoothingFilter.Type.Linear if linear else pv.SmoothingFilter.Type.Const)
    sf.setInputDimensions(list(dims))
    assert rel(sf.apply(x), orc.smoothing_filter(x, dims, radius + 1.0, linear)) < 1e-13
    sf.rmin = 2.5
    assert rel(sf.apply(x), orc.smoothing_filter(x, dims, 2.5, linear)) < 1e-13


def test_projection_and_chain(pv):
    dims = (10, 6, 8)
    n = int(np.prod(dims))
    x = RNG.uniform(0.2, 0.4, n)
    sf = pv.SmoothingFilter(2, pv.SmoothingFilter.Type.Linear)
    pf = pv.ProjectionFilter(8.0)
    chain = pv.FilterChain([sf, pf], list(dims))
    chain.setDesignVars(x)
    s1 = orc.smoothing_filter(x, dims, 3.0, True)
    p1 = orc.projection(s1, 8.0)
    assert rel(chain.physicalVars(), p1) < 1e-13
    g = RNG.normal(size=n)
    ref = orc.smoothing_filter(orc.projection_backprop(g, s1, 8.0), dims, 3.0, True)
    assert rel(chain.backprop(g), ref) < 1e-12
    assert pf.invert(0.3) == pytest.approx(np.arctanh((2 * 0.3 - 1) * np.tanh(4.0)) / 8 + 0.5, rel=1e-14)


def test_layer_by_layer_zero_init(pv):
    ne = (16, 8, 8)
    rho = wl.noise_density(ne, 3)
    full = wl.configure(pv, ne, None, rho)
    lbl_sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = lbl_sim.multigridSolver(2)
    ev = pv.LayerByLayerEvaluator(lbl_sim)
    ev.selectInitMethod("zero")
    ev.run(mg, True, 2, 500, 1e-10)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    X = wl._node_positions(ne)
    s.set_dirichlet(np.where(X[:, 1] == 0, 7, 0).astype(np.uint8))
    s.set_gravity([0.0, -1.0, 0.0])
    om = orc.MG(s, 2)
    obj, grad, comp, its = om.lbl_run_zero(2, 500, 1e-10)
    assert ev.objective() == pytest.approx(obj, rel=1e-9)
    assert rel(ev.gradient(), grad) < 1e-8


def test_layer_by_layer_initial_guesses(pv):
    """The reference's initial guesses (LayerByLayer.hh:24-221): every guess converges to the same layer
    solutions, so objective and gradient match the zero-guess oracle; the subspace guess (the default,
    "N=3") needs fewer PCG iterations than zero; a second run with zeroInit=False starts the full design
    from the previous run's full-design solution, which already meets the tolerance."""
    ne = (16, 8, 8)
    rho = wl.noise_density(ne, 4)
    full = wl.configure(pv, ne, None, rho)
    lbl_sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = lbl_sim.multigridSolver(2)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    X = wl._node_positions(ne)
    s.set_dirichlet(np.where(X[:, 1] == 0, 7, 0).astype(np.uint8))
    s.set_gravity([0.0, -1.0, 0.0])
    obj, grad, comp, its_ref = orc.MG(s, 2).lbl_run_zero(1, 500, 1e-12)
    iters = {}
    for method in ("zero", "constant", "fd", "N=1", "N=3"):
        ev = pv.LayerByLayerEvaluator(lbl_sim)
        if method != "N=3":  # "N=3" is the default
            ev.selectInitMethod(method)
        ev.batched = False
        ev.run(mg, True, 1, 500, 1e-12)
        assert ev.objective() == pytest.approx(obj, rel=1e-10), method
        assert rel(ev.gradient(), grad) < 1e-9, method
        assert np.allclose(ev.perLayerCompliance(), comp, rtol=1e-9, atol=0), method
        iters[method] = sum(ev.perLayerIterations())
        if method == "N=3":
            ev.run(mg, False, 1, 500, 1e-12)  # warm start of the full design
            assert ev.perLayerIterations()[0] <= 1
            assert ev.objective() == pytest.approx(obj, rel=1e-10)
    assert iters["N=3"] < iters["zero"]
    with pytest.raises(Exception, match="Unrecognized method"):
        pv.LayerByLayerEvaluator(lbl_sim).selectInitMethod("N=x")


# ---------------------------------------------------------------- RAMP law (TensorProductSimulator.hh:2055-2060,
# 997-1000): the law the reference's layer-by-layer drivers use by default (q = 3)
def _ramp_pair(pv, ne, q=3.0, seed=5):
    tps, s = make_pair(pv, ne, seed=seed)
    tps.interpolationLaw = pv.InterpolationLaw.RAMP
    tps.q = q
    s.set_law(1, 1.0, 1e-9, 3.0, q)
    return tps, s


def test_ramp_young_and_gradient(pv):
    ne = (16, 8, 8)
    tps, s = _ramp_pair(pv, ne)
    assert tps.interpolationLaw == pv.InterpolationLaw.RAMP and tps.q == 3.0
    assert rel(tps.getYoungModulusScaleFactor(), s.E()) < 1e-14
    u = RNG.normal(size=(tps.numNodes(), 3))
    assert rel(tps.complianceGradient(u), s.compliance_gradient(u)) < 1e-12
    tps.gravity = [0.0, -1.0, 0.0]
    s.set_gravity([0.0, -1.0, 0.0])
    assert rel(tps.buildLoadVector(), s.build_load()) < 1e-14
    assert rel(tps.complianceGradient(u), s.compliance_gradient(u)) < 1e-12
    # masked layers: moduli of masked elements are zero under either law
    tps.setFabricationMaskHeightByLayer(5)
    om = orc.MG(s, 1)
    om.set_mask_layer(5)
    assert rel(tps.getYoungModulusScaleFactor(), s.E()) < 1e-14
    assert rel(tps.applyK(u), s.apply_K(u)) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ramp_pcg_matches_oracle(pv, precision):
    ne = (32, 16, 16)
    tps, s = _ramp_pair(pv, ne)
    mg, om = tps.multigridSolver(2), orc.MG(s, 2)
    f = tps.buildLoadVector()
    u0 = np.zeros_like(f)
    u = mg.preconditionedConjugateGradient(u0, f, 500, 1e-8, precision=precision)
    ref, _ = om.pcg(u0, f, 500, 1e-8)
    c, cref = 0.5 * np.sum(f * u), 0.5 * np.sum(f * ref)
    if precision == "fp64":
        assert rel(u, ref) < 1e-10
        assert abs(c - cref) <= 1e-12 * abs(cref)
    else:
        assert abs(c - cref) <= 1e-9 * abs(cref)
    g, gref = tps.complianceGradient(u), s.compliance_gradient(ref)
    assert rel(g, gref) < (1e-10 if precision == "fp64" else 1e-7)


@pytest.mark.parametrize("method,batched", [("zero", True), ("zero", False), ("N=3", False)])
def test_layer_by_layer_ramp(pv, method, batched):
    """Layer-by-layer evaluation under RAMP: getIntermediateFabricationShape(..., law=RAMP) gives the
    intermediate simulator the default RAMP factor q = 3 (TensorProductSimulator.hh:1876, 2138)."""
    ne = (16, 8, 8)
    rho = wl.noise_density(ne, 6)
    full = wl.configure(pv, ne, None, rho)
    full.q = 7.0  # not carried over to the intermediate shape
    lbl_sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.RAMP)
    assert lbl_sim.interpolationLaw == pv.InterpolationLaw.RAMP and lbl_sim.q == 3.0
    mg = lbl_sim.multigridSolver(2)
    ev = pv.LayerByLayerEvaluator(lbl_sim)
    ev.selectInitMethod(method)
    ev.batched = batched
    ev.run(mg, True, 2, 500, 1e-11)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(1, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    X = wl._node_positions(ne)
    s.set_dirichlet(np.where(X[:, 1] == 0, 7, 0).astype(np.uint8))
    s.set_gravity([0.0, -1.0, 0.0])
    obj, grad, comp, its = orc.MG(s, 2).lbl_run_zero(2, 500, 1e-11)
    assert ev.objective() == pytest.approx(obj, rel=1e-9)
    assert rel(ev.gradient(), grad) < 1e-8
    assert np.allclose(ev.perLayerCompliance(), comp, rtol=1e-9, atol=0)


# ---------------------------------------------------------------- OC update with non-finite ratios
def test_oc_update_non_finite_ratios_match_oracle():
    """OptimalityCriterion.hh:64-75 through vf_oc_update on device data: a NaN ratio (0/0, a positive
    sensitivity under p = 0.5, 0 * inf) keeps x0, an infinite one is clamped to x0 + m."""
    import ctypes

    import torch

    from voxelfem_amd import capi

    n = 4096
    x0 = RNG.uniform(0.0, 1.0, n)
    dJ = -RNG.uniform(0.0, 2.0, n)
    dc = np.full(n, -1.0 / n)
    dJ[::7] = -dJ[::7]          # positive sensitivities: NaN ratios
    dJ[3::11] = 0.0
    dc[3::11] = 0.0             # 0/0
    dc[5::13] = 0.0             # -1/0 = -inf, (-inf)^0.5 = inf
    x0[5::26] = 0.0             # 0 * inf = NaN
    dev = torch.device("cuda:0")
    tx, tJ, tc = (torch.tensor(a, device=dev) for a in (x0, dJ, dc))
    out = torch.empty_like(tx)
    L = capi.lib()
    L.vf_oc_update.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    capi.check(L.vf_oc_update(n, tx.data_ptr(), tJ.data_ptr(), tc.data_ptr(), 1.7, 0.2, 0.5, out.data_ptr(),
                              capi.VF_DEVICE, None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = orc.oc_update(x0, dJ, dc, 1.7, 0.2, 0.5)
    with np.errstate(all="ignore"):
        raw = x0 * (dJ / (dc * 1.7)) ** 0.5
    special = ~np.isfinite(raw)
    # non-finite ratios: the kept x0 (NaN) or the clamped bound (inf), exactly
    bad = np.flatnonzero(got[special] != ref[special])
    print(f"{special.sum()} non-finite ratios, {bad.size} differ; finite entries max rel diff "
          f"{np.max(np.abs(got[~special] - ref[~special]) / np.abs(ref[~special])):.2e}")
    assert bad.size == 0, (x0[special][bad[:5]], got[special][bad[:5]], ref[special][bad[:5]])
    # NaN ratios in C semantics (numpy's power gives NaN for (-inf)^0.5 where C's pow gives +inf): 0/0, a
    # negative finite base, and 0 * inf; each keeps x0
    with np.errstate(all="ignore"):
        q = dJ / (dc * 1.7)
    nan_ratio = np.isnan(q) | (np.isfinite(q) & (q < 0)) | ((x0 == 0) & np.isinf(q))
    assert nan_ratio.sum() > 100
    assert np.all(got[nan_ratio] == x0[nan_ratio])
    # finite ratios: device pow and the host's may differ in the last bit
    np.testing.assert_allclose(got[~special], ref[~special], rtol=1e-14, atol=0)


def test_oc_step_self_weight_matches_oracle(pv):
    """One OC step of a self-weight problem (positive sensitivities where gravity dominates, i.e. NaN
    ratios) through TopologyOptimizationProblem + OCOptimizer, against the oracle's OC step."""
    import voxelfem_amd

    ne = (16, 8, 8)
    n = int(np.prod(ne))
    vfrac = 0.3
    x0 = wl.noise_density(ne, 8)
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, x0)
    tps.gravity = [0.0, 0.0, -0.05]
    levels = voxelfem_amd.optMGLevel(ne, 3)
    mg = tps.multigridSolver(levels)
    obj = pv.MultigridComplianceObjective(mg)
    obj.tol = 1e-12
    obj.cgIter = 1000
    top = pv.TopologyOptimizationProblem(tps, obj, [pv.TotalVolumeConstraint(vfrac)], [])
    top.setVars(x0)
    g = top.evaluateObjectiveGradient()
    assert np.any(g > 0), "the test needs positive sensitivities"
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(x0)
    s.set_dirichlet(wl.cantilever_mask(ne))
    s.set_forces(wl.cantilever_load(ne))
    s.set_gravity([0.0, 0.0, -0.05])
    f = s.build_load()
    u, _ = orc.MG(s, levels).pcg(np.zeros_like(f), f, 1000, 1e-12, 1, 2, True)
    gref = s.compliance_gradient(u)
    assert np.abs(g - gref).max() <= 1e-9 * np.abs(gref).max()
    dc = np.full(n, -1.0 / (vfrac * n))
    assert np.allclose(top.evaluateConstraintsJacobian()[0], dc, rtol=1e-12, atol=0)
    pv.OCOptimizer(top).step()
    xr = orc.OC().step(x0, gref, dc, vfrac, orc.Chain(ne))
    got = top.getVars()
    assert np.abs(got - xr).max() <= 1e-9
    keep = gref > 0
    assert np.all(got[keep] == x0[keep])
