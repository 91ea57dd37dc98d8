"""GPU tests of the LDS-tiled fine-level kernels (structured element matrix, two-launch smoother) against
the register-gather kernels (one launch per colour) and against the oracle.

Grids are chosen so the (y, z) tiles split unevenly (node counts 39 and 71 against tiles of 16 and 32),
with fabrication mask layers (detached rows that the smoother must leave untouched) and with partially
constrained nodes (the MBB roller and symmetry plane) that take the sequential point update.

The element-centric Walsh-basis K u (variant 3) is checked the same way. Tolerances: K u and one
smoothing sweep agree to 1e-13 relative between kernel families (they round in a different order); a
V-cycle in FP64 to 1e-12; in FP32 to 1e-5 (single-precision rounding).
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from voxelfem_amd import workloads as wl

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(1234)


@pytest.fixture(scope="module")
def vf():
    import voxelfem_amd

    yield voxelfem_amd
    _variant(0)


def _variant(v):
    from voxelfem_amd import capi

    L = capi.lib()
    L.vf_set_kernel_variant.argtypes = [ctypes.c_int]
    capi.check(L.vf_set_kernel_variant(v))


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def both(fn, variant=2):
    """fn() evaluated with the tiled (2) or Walsh (3) fine kernels and with the register-gather ones (1)."""
    _variant(variant)
    a = fn()
    _variant(1)
    try:
        b = fn()
    finally:
        _variant(0)
    return a, b


CASES = [
    ((10, 38, 70), wl.CANTILEVER_BC, False, None),
    ((10, 38, 70), wl.MBB_BC, True, None),
    ((12, 38, 22), wl.CANTILEVER_BC, False, 17),
    ((8, 20, 40), wl.MBB_BC, True, 9),
]


@pytest.mark.parametrize("ne,bc,sym,layer", CASES)
def test_tiled_fine_operators_match_gather(vf, ne, bc, sym, layer):
    pv = vf.pyVoxelFEM
    rho = wl.noise_density(ne, 3)
    tps = wl.configure(pv, ne, bc, rho, symmetry_x=sym)
    mg = tps.multigridSolver(1)
    if layer is not None:
        tps.setFabricationMaskHeightByLayer(layer)
        mg.setFabricationMaskHeightByLayer(layer)
    nn = tps.numNodes()
    u = RNG.normal(size=(nn, 3))
    b = RNG.normal(size=(nn, 3))
    for var in (2, 3):
        ku_t, ku_g = both(lambda: tps.applyK(u), var)
        assert rel(ku_t, ku_g) < 1e-13, var
        r_t, r_g = both(lambda: mg.computeResidual(0, u, b), var)
        assert rel(r_t, r_g) < 1e-13, var
    for fwd in (True, False):
        for var in (2, 4):
            s_t, s_g = both(lambda: mg._smooth(0, u, b, fwd), var)
            assert rel(s_t, s_g) < 1e-13, (fwd, var)
        if layer is not None:
            # rows of detached nodes keep their input values
            ny = ne[1] + 1
            iy = (np.arange(nn) // (ne[2] + 1)) % ny
            det = iy >= layer + 1
            assert np.array_equal(s_t[det], u[det])


@pytest.mark.parametrize("ne,bc,sym,layer", CASES[:2])
def test_tiled_smoother_matches_oracle(vf, ne, bc, sym, layer):
    pv = vf.pyVoxelFEM
    rho = wl.noise_density(ne, 4)
    tps = wl.configure(pv, ne, bc, rho, symmetry_x=sym)
    mg = tps.multigridSolver(1)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    m = tps.getDirichletMask().astype(np.uint8)
    s.set_dirichlet((m[:, 0] | (m[:, 1] << 1) | (m[:, 2] << 2)).astype(np.uint8))
    om = orc.MG(s, 1)
    nn = tps.numNodes()
    u = RNG.normal(size=(nn, 3))
    b = RNG.normal(size=(nn, 3))
    for fwd in (True, False):
        assert rel(mg._smooth(0, u, b, fwd), om.smooth(0, u, b, fwd)) < 1e-12, fwd


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("variant", [2, 3, 4])
def test_tiled_vcycle_matches_gather(vf, prec, tol, variant):
    pv = vf.pyVoxelFEM
    ne = (16, 24, 40)
    tps = wl.configure(pv, ne, wl.MBB_BC, wl.noise_density(ne, 5), symmetry_x=True)
    mg = tps.multigridSolver(3)
    r = RNG.normal(size=(tps.numNodes(), 3))
    for nu in (1, 2):
        a, b = both(lambda: mg._precondition(r, 1, nu, False, prec), variant)
        assert rel(a, b) < tol, nu


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_level1_matrix_cores_vcycle(vf, prec, tol):
    """A V-cycle whose level 1 runs on the matrix cores (variant 5: MFMA residual and smoother from the
    fine moduli) against level 1 on the vector units (variant 1), on a grid whose level-1 colour classes
    do not fill the 16-node MFMA tiles, with the MBB roller and symmetry plane constraints."""
    pv = vf.pyVoxelFEM
    ne = (16, 24, 40)
    tps = wl.configure(pv, ne, wl.MBB_BC, wl.noise_density(ne, 6), symmetry_x=True)
    mg = tps.multigridSolver(3)
    r = RNG.normal(size=(tps.numNodes(), 3))

    def run(nu):
        mg.updateStiffnessMatrices()  # the level-1 form is fixed when the hierarchy is built
        return mg._precondition(r, 1, nu, False, prec)

    for nu in (1, 2):
        a, b = both(lambda: run(nu), 5)
        assert rel(a, b) < tol, nu


# ----------------------------------------------------------------------------- graph-replayed PCG
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pcg_graph_replay_matches_eager(vf, prec):
    """The steady-state PCG iteration replayed from a captured HIP graph gives the eager iterates: the
    same solve, run eagerly, with the graph, again eagerly and again with a fresh capture, agrees bit for
    bit (each replay is checked through the final iterate and the residual history)."""
    pv = vf.pyVoxelFEM
    ne = (32, 16, 16)
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, wl.noise_density(ne, 6))
    mg = tps.multigridSolver(2)
    f = tps.buildLoadVector()

    def solve(graphs):
        mg.useGraphs = graphs
        hist = []
        u = mg.preconditionedConjugateGradient(np.zeros_like(f), f, 200, 1e-10,
                                               lambda it, x, r: hist.append(float(np.abs(r).max())),
                                               precision=prec)
        return u, hist

    u0, h0 = solve(False)
    u1, h1 = solve(True)
    u2, h2 = solve(False)
    u3, h3 = solve(True)
    assert len(h0) > 3
    for u, h in ((u1, h1), (u2, h2), (u3, h3)):
        assert np.array_equal(u, u0)
        assert h == h0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pcg_kernel_timing_in_replayed_graph(vf, prec):
    """Timing the fine K d of the fused PCG where it runs (events on the solver's stream, captured into the
    replayed graph) times one launch per iteration and leaves the iterates unchanged."""
    import torch

    from voxelfem_amd import capi

    pv = vf.pyVoxelFEM
    ne = (32, 16, 16)
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, wl.noise_density(ne, 8))
    mg = tps.multigridSolver(2)
    sim_h, mg_h = tps._handle(), mg._handle()
    nn = tps.numNodes()
    dev = torch.device("cuda:0")
    f = torch.zeros(3 * nn, dtype=torch.float64, device=dev)
    capi.check(capi.lib().vf_sim_build_load(sim_h, f.data_ptr(), capi.VF_DEVICE))
    p = capi.VF_PREC_FP64 if prec == "fp64" else capi.VF_PREC_FP32
    out = []
    for graphs, timing in ((False, False), (True, True), (False, True)):
        mg.useGraphs = graphs
        capi.set_kernel_timing(mg_h, timing)
        x = torch.zeros_like(f)
        st = capi.pcg(mg_h, x.data_ptr(), f.data_ptr(), 500, 1e-10, 1, 1, False, p)
        torch.cuda.synchronize()
        n, ms = capi.kernel_timing(mg_h)
        capi.set_kernel_timing(mg_h, False)
        out.append(x.cpu().numpy())
        assert n == (st.iterations if timing else 0)
        assert (ms > 0) == timing
    mg.useGraphs = False
    assert np.array_equal(out[1], out[0]) and np.array_equal(out[2], out[0])


# ----------------------------------------------------------------------------- batched layer-by-layer
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lbl_batched_matches_sequential_and_oracle(vf, prec):
    """The layer solves spread over concurrent pipelines with graph-replayed PCG iterations give the
    sequential evaluator's objective, gradient and per-layer iteration counts (results are summed in
    layer order), and the oracle's objective and gradient."""
    pv = vf.pyVoxelFEM
    ne = (16, 12, 8)
    rho = wl.noise_density(ne, 7)
    full = wl.configure(pv, ne, None, rho)
    lbl_sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = lbl_sim.multigridSolver(2)
    ev = pv.LayerByLayerEvaluator(lbl_sim)
    ev.selectInitMethod("zero")
    ev.precision = prec
    ev.batched = False
    ev.run(mg, True, 1, 500, 1e-10)
    obj_seq, g_seq, its_seq = ev.objective(), ev.gradient(), list(ev.perLayerIterations())
    for pipes, graphs in ((3, True), (1, True), (5, False)):
        ev.batched, ev.pipelines, ev.useGraphs = True, pipes, graphs
        ev.run(mg, True, 1, 500, 1e-10)
        assert ev.objective() == pytest.approx(obj_seq, rel=1e-13), (pipes, graphs)
        assert rel(ev.gradient(), g_seq) < 1e-13, (pipes, graphs)
        assert list(ev.perLayerIterations()) == its_seq, (pipes, graphs)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    X = wl._node_positions(ne)
    s.set_dirichlet(np.where(X[:, 1] == 0, 7, 0).astype(np.uint8))
    s.set_gravity([0.0, -1.0, 0.0])
    om = orc.MG(s, 2)
    obj, grad, comp, its = om.lbl_run_zero(1, 500, 1e-10)
    tol = 1e-9 if prec == "fp64" else 1e-8
    assert ev.objective() == pytest.approx(obj, rel=tol)
    assert rel(ev.gradient(), grad) < 10 * tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lbl_batched_on_reused_nan_memory(vf, prec):
    """Regression: the layer solves read detached rows of the PCG direction vector (times zero moduli), so
    those rows must be initialised even when freshly allocated device memory holds NaN bit patterns. The
    device memory the pipelines will allocate is first filled with NaN and handed back to the runtime."""
    import torch

    pv = vf.pyVoxelFEM
    ne = (16, 12, 8)
    rho = wl.noise_density(ne, 7)
    full = wl.configure(pv, ne, None, rho)
    lbl_sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = lbl_sim.multigridSolver(2)
    ev = pv.LayerByLayerEvaluator(lbl_sim)
    ev.selectInitMethod("zero")
    ev.precision = prec
    ev.batched = False
    ev.run(mg, True, 1, 500, 1e-10)
    obj_seq = ev.objective()
    junk = torch.full((1 << 26,), float("nan"), dtype=torch.float64, device="cuda:0")  # 512 MiB
    torch.cuda.synchronize()
    del junk
    torch.cuda.empty_cache()
    ev.batched, ev.pipelines, ev.useGraphs = True, 3, True
    ev.run(mg, True, 1, 500, 1e-10)
    assert np.isfinite(ev.objective())
    assert ev.objective() == pytest.approx(obj_seq, rel=1e-13)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_zero_start_vcycle_masked_layers_on_nan_memory(vf, prec, tol):
    """The V-cycle from a zero fine iterate with the two-launch smoother (variant 2) clears only the detached
    rows of the iterate and lets its first sweep read the rest as zero without loading it. On a structure with
    masked top layers, and with the device memory the new solver allocates first filled with NaN and handed
    back to the runtime, the cycle is finite and equals the register-gather family's (variant 1) and, in
    FP64, the oracle's."""
    import torch

    pv = vf.pyVoxelFEM
    ne = (16, 24, 40)
    rho = wl.noise_density(ne, 8)
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, rho)
    layer = 17
    tps.setFabricationMaskHeightByLayer(layer)
    junk = torch.full((1 << 26,), float("nan"), dtype=torch.float64, device="cuda:0")  # 512 MiB
    torch.cuda.synchronize()
    del junk
    torch.cuda.empty_cache()
    mg = tps.multigridSolver(3)
    r = RNG.normal(size=(tps.numNodes(), 3))
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(rho)
    s.set_dirichlet(wl.cantilever_mask(ne))
    om = orc.MG(s, 3)
    om.set_mask_layer(layer)
    om.update_stiffness()
    for nu in (1, 2):
        a, b = both(lambda: mg._precondition(r, 1, nu, False, prec), 2)
        assert np.isfinite(a).all() and np.isfinite(b).all(), nu
        assert rel(a, b) < tol, nu
        if prec == "fp64":
            assert rel(a, om.precondition(r, 1, nu)) < 1e-12, nu
