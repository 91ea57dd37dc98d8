"""The rest of the reference's pyVoxelFEM surface (python_bindings/VoxelFEM.cc:76-193) through the HIP library:
getK, constantStrainLoad, sampleNodalField, the intermediate-shape transfers, getMesh, debug_get_x/b, the
element visit order, the ETensor / dx properties and the layer-by-layer callbacks, each against the oracle's
restatement or its definition."""
import numpy as np
import pytest

from oracle import oracle as orc
from voxelfem_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pv():
    import voxelfem_amd

    return voxelfem_amd.pyVoxelFEM


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _sim(pv, ne, lo, hi, seed=0):
    tps = pv.TensorProductSimulator([1, 1, 1], [list(lo), list(hi)], list(ne))
    tps.ETensor = (1.3, 0.3)
    tps.E_min = 1e-9
    rho = np.random.default_rng(seed).uniform(0.05, 1.0, int(np.prod(ne)))
    tps.setDensities(rho)
    return tps, rho


def test_getK_upper_triangle_matches_oracle_and_applyK(pv):
    ne, lo, hi = (4, 3, 2), (0, 0, 0), (2.0, 1.5, 1.0)
    tps, _ = _sim(pv, ne, lo, hi)
    K = tps.getK()
    ref = orc.stiffness_upper(ne, tps.fullDensityElementStiffnessMatrix(), tps.getYoungModulusScaleFactor())
    assert K.shape == ref.shape
    assert abs(K - ref).max() <= 1e-14 * abs(ref).max()
    u = np.random.default_rng(1).normal(size=(tps.numNodes(), 3))
    Kf = K + K.T - np.diag(K.diagonal())
    assert rel((Kf @ u.reshape(-1)).reshape(-1, 3), tps.applyK(u)) < 1e-13


def test_constant_strain_load_matches_oracle(pv):
    ne, lo, hi = (6, 4, 5), (0.5, -1, 0), (3.5, 1.0, 2.5)
    tps, rho = _sim(pv, ne, lo, hi, seed=2)
    A = np.random.default_rng(3).normal(size=(3, 3))
    eps = 0.5 * (A + A.T)
    got = tps.constantStrainLoad(eps)
    h = (np.asarray(hi, float) - np.asarray(lo, float)) / np.asarray(ne)
    ref = orc.constant_strain_load(ne, h, 1.3, 0.3, rho, eps)
    assert rel(got, ref) < 1e-14
    six = np.array([eps[0, 0], eps[1, 1], eps[2, 2], eps[1, 2], eps[0, 2], eps[0, 1]])
    assert np.array_equal(tps.constantStrainLoad(six), got)
    with pytest.raises(RuntimeError, match="symmetric"):
        tps.constantStrainLoad(A + np.triu(np.ones((3, 3)), 1))


def test_sample_nodal_field_matches_oracle(pv):
    ne, lo, hi = (5, 4, 3), (0.5, -1.0, 0.0), (3.0, 1.0, 1.5)
    tps, _ = _sim(pv, ne, lo, hi)
    rng = np.random.default_rng(4)
    u = rng.normal(size=(tps.numNodes(), 3))
    lo_, hi_ = np.asarray(lo, float), np.asarray(hi, float)
    pts = np.concatenate([rng.uniform(lo_ - 0.2, hi_ + 0.2, size=(300, 3)), hi_[None], lo_[None]])
    got = tps.sampleNodalField(u, pts)
    ref = orc.sample_nodal_field(ne, lo, hi, u, pts)
    assert got.shape == (len(pts), 3)
    assert rel(got, ref) < 1e-13
    s1 = tps.sampleNodalField(u[:, :1].copy(), pts)
    assert s1.shape == (len(pts), 1) and rel(s1[:, 0], ref[:, 0]) < 1e-13


def test_intermediate_shape_transfers(pv):
    ne = (6, 8, 4)
    tps, rho = _sim(pv, ne, (0, 0, 0), ne)
    part = tps.getIntermediateFabricationShape(0.5, False, pv.InterpolationLaw.SIMP)
    pne = part.NbElementsPerDimension
    assert tuple(pne) == (6, 4, 4)
    R = rho.reshape(ne)
    part.setUniformDensities(0.0)
    tps.transferDensitiesToIntermediateFabricationShape(part)
    assert np.array_equal(part.getDensities(), R[:, :4, :].ravel())
    u = np.random.default_rng(5).normal(size=(tps.numNodes(), 3))
    U = u.reshape(7, 9, 5, 3)
    assert np.array_equal(tps.transferVFieldToIntermediateFabricationShape(part, u), U[:, :5, :, :].reshape(-1, 3))
    acc = np.ones(tps.numElements())
    rin = np.random.default_rng(6).normal(size=part.numElements())
    tps.accumElementScalarFieldFromIntermediateFabricationShape(part, rin, acc)
    exp = np.ones(ne)
    exp[:, :4, :] += rin.reshape(pne)
    assert np.array_equal(acc, exp.ravel())


def test_get_mesh_matches_oracle(pv):
    ne, lo, hi = (3, 2, 4), (0.5, 0, -1), (2.0, 1.0, 1.0)
    tps, _ = _sim(pv, ne, lo, hi)
    V, F = tps.getMesh()
    Vr, Fr = orc.mesh(ne, lo, hi)
    assert np.allclose(V, Vr, rtol=0, atol=1e-15)
    assert np.array_equal(F, Fr)
    assert tps.mesh() is tps


def test_element_visit_order_is_colour_classes(pv):
    ne = (4, 3, 5)
    tps, _ = _sim(pv, ne, (0, 0, 0), ne)
    order = tps.debugMulticolorElementVisit()
    assert sorted(order) == list(range(int(np.prod(ne))))
    E = order.reshape(ne)
    # elements of one colour (index parities) come as a block; no two of them share a node
    colour = (np.indices(ne) % 2)
    cid = 4 * colour[0] + 2 * colour[1] + colour[2]
    for c in range(8):
        vals = np.sort(E[cid == c])
        assert np.array_equal(vals, np.arange(vals[0], vals[0] + vals.size))


def test_etensor_object_and_dx(pv):
    ne = (4, 2, 2)
    tps, _ = _sim(pv, ne, (0, 0, 0), (4.0, 1.0, 1.0))
    e = tps.ETensor
    assert (e.E, e.nu) == (1.3, 0.3) and (e[0], e[1]) == (1.3, 0.3)
    e.setIsotropic(1.0, 0.25)
    tps.ETensor = e
    assert np.allclose(tps.fullDensityElementStiffnessMatrix(), orc.element_stiffness(1.0, 0.25, [1, 0.5, 0.5]),
                       rtol=0, atol=1e-15)
    tps.dx = [0.5, 0.5, 0.25]
    assert np.allclose(tps.dx, [0.5, 0.5, 0.25]) and tps.elementVolume == pytest.approx(0.0625)
    assert np.allclose(tps.fullDensityElementStiffnessMatrix(), orc.element_stiffness(1.0, 0.25, [0.5, 0.5, 0.25]),
                       rtol=0, atol=1e-15)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_debug_get_x_b(pv, precision):
    """debug_get_x/b(l) (MultigridSolver.hh:1158-1159): level 0 holds the last cycle's fine iterate and right-hand
    side, the coarsest level its direct solve (K_L x_L = b_L on the free components)."""
    ne = (16, 8, 8)
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, wl.noise_density(ne, 1))
    mg = tps.multigridSolver(2)
    f = tps.buildLoadVector()
    with pytest.raises(RuntimeError, match="no multigrid cycle"):
        mg.debug_get_x(0)
    if precision == "fp64":
        u = mg.solve(np.zeros_like(f), f, 1, 1)
        assert np.array_equal(mg.debug_get_x(0), u)
        assert np.array_equal(mg.debug_get_b(0), f)
    else:
        s = mg._precondition(f, 1, 1, False, "fp32")
        assert rel(mg.debug_get_x(0), s) < 1e-6
    L = 2
    xL, bL = mg.debug_get_x(L), mg.debug_get_b(L)
    r = mg.zeroOutDirichletComponents(L, mg.applyK(L, xL) - bL)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert np.abs(r).max() <= tol * np.abs(bL).max()
    assert np.abs(mg.debug_get_b(1)).max() > 0


def test_layer_by_layer_callbacks(pv):
    """it_callback and lblCallback (LayerByLayer.hh:222-279): the callback's per-layer compliance and gradient
    sum to the objective and gradient, its u solves the layer system, and the PCG callback sees every iteration."""
    ne = (16, 8, 8)
    rho = wl.noise_density(ne, 3)
    full = wl.configure(pv, ne, None, rho)
    sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = sim.multigridSolver(2)
    seen, its = [], [0]

    def lbl_cb(l, compliance, grad, u):
        seen.append((l, compliance, grad.copy(), u.copy()))

    def it_cb(i, x, r):
        its[0] += 1

    for method in ("zero", "N=3"):
        seen.clear()
        its[0] = 0
        ev = pv.LayerByLayerEvaluator(sim)
        ev.selectInitMethod(method)
        ev.run(mg, True, 2, 500, 1e-10, it_callback=it_cb, lblCallback=lbl_cb)
        assert [s[0] for s in seen] == list(range(8, 0, -2))
        assert np.allclose([s[1] for s in seen], ev.perLayerCompliance(), rtol=0, atol=0)
        assert ev.objective() == pytest.approx(0.5 * sum(s[1] for s in seen) / len(seen), rel=1e-14)
        assert rel(sum(s[2] for s in seen) / len(seen), ev.gradient()) < 1e-13
        assert its[0] >= sum(ev.perLayerIterations())
        # the last callback's u is the solution of the lowest layer (mask at layer 2)
        l, c, g, u = seen[-1]
        assert c == pytest.approx(float(np.sum(sim.buildLoadVector() * u)), rel=1e-12)


@pytest.mark.parametrize("method", ["zero", "N=3"])
def test_partial_band_update_equals_full_rebuild(pv, method):
    """The layer-by-layer evaluator updates the hierarchy by bands after each layer removal
    (MultigridSolver.hh:907-1017). The recomputed rows equal a full rebuild's, so every layer's compliance, the
    iteration counts and the gradient are identical with and without partial updates (one solver, so both runs
    use the kernels its autotuner chose)."""
    ne = (32, 16, 16)
    rho = wl.noise_density(ne, 9)
    full = wl.configure(pv, ne, None, rho)
    sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = sim.multigridSolver(3)
    ev = pv.LayerByLayerEvaluator(sim)
    ev.selectInitMethod(method)
    ev.batched = False
    count = 0
    for prec in ("fp64", "fp32"):
        ev.precision = prec
        out = {}
        for partial in (False, True):
            mg.partialStiffnessUpdates = partial
            ev.run(mg, True, 3, 500, 1e-10)
            out[partial] = (ev.objective(), ev.gradient(), ev.perLayerCompliance(), ev.perLayerIterations())
            nl = len(ev.perLayerCompliance())
            count += (nl - 1) if partial else 0
            assert mg.partialUpdateCount == count
        a, b = out[False], out[True]
        print(f"{method} {prec}: objective {a[0]!r} vs {b[0]!r}, {nl} layers")
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1])
        assert a[2] == b[2] and a[3] == b[3]


def test_mma_drives_a_topology_problem(pv):
    """pyOptimizer.MMA as python/LayerByLayerOptimization.py:13-40 drives it: objective and volume constraint
    from a TopologyOptimizationProblem on the device, gradients stacked (m + 1) x n. The compliance decreases and
    the volume constraint is met."""
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "voxelfem_amd"))
    try:
        import pyOptimizer
    finally:
        sys.path.pop(0)
    ne = (16, 8, 8)
    n = int(np.prod(ne))
    vfrac = 0.3
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, np.full(n, vfrac))
    mg = tps.multigridSolver(2)
    obj = pv.MultigridComplianceObjective(mg)
    top = pv.TopologyOptimizationProblem(tps, obj, [pv.TotalVolumeConstraint(vfrac)],
                                         [pv.SmoothingFilter(1, pv.SmoothingFilter.Type.Linear)])

    def objAndConstr(x):
        top.setVars(x)
        return np.array([top.evaluateObjective(), -top.evaluateConstraints()[0]])

    def gradients(x):
        return np.stack([top.evaluateObjectiveGradient(), -top.evaluateConstraintsJacobian()[0]])

    opt = pyOptimizer.MMA(n, 1, np.zeros(n), np.ones(n), objAndConstr, gradients)
    opt.setInitialVar(np.full(n, vfrac))
    top.setVars(np.full(n, vfrac))
    c0 = top.evaluateObjective()
    for _ in range(8):
        opt.step()
    x = opt.getOptimalVar()
    top.setVars(x)
    c1 = top.evaluateObjective()
    print(f"MMA: compliance {c0:.6g} -> {c1:.6g}, volume constraint {top.evaluateConstraints()[0]:.2e}")
    assert c1 < 0.8 * c0
    assert top.evaluateConstraints()[0] >= -1e-3
    assert x.min() >= 0 and x.max() <= 1
