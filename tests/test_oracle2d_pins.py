"""Pins the planar (2D) CPU oracle (oracle/vf_oracle2d.cpp) against what the mathematics fixes.

The reference ships no 2D golden vectors, and its isotropic tensor comes from an un-vendored submodule
(MeshFEM's ElasticityTensor), so every check here is a closed form (the published plane-stress Q1 matrix of
the unit square, constant-strain energies), an invariant, or a brute-force sparse construction on a tiny grid
(tests/fem_reference2d.py), chosen so that a dropped term, a wrong sign or index, a plane-strain tensor or a
transposed operand fails.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import oracle2d as o2
from tests import fem_reference2d as fr

RNG = np.random.default_rng(4321)


def make_sim(ne, lo=None, hi=None, E=1.0, nu=0.3, Emin=1e-9, rho=None):
    s = o2.Sim(ne, lo, hi)
    s.set_etensor(E, nu)
    s.set_law(0, 1.0, Emin, 3.0, 3.0)
    if rho is None:
        rho = RNG.uniform(0.2, 1.0, s.nelems)
    s.set_densities(rho)
    return s


# ------------------------------------------------------------------------- K0 (TensorProductSimulator.hh:67-80)
@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_K0_equals_published_plane_stress_matrix(nu):
    KE = fr.top88_KE(1.0, nu)
    K0 = o2.element_stiffness(1.0, nu, (1.0, 1.0))
    perm = [0, 2, 3, 1]  # (0,0), (1,0), (1,1), (0,1) in local numbering l = 2 i + j
    dof = np.array([[2 * perm[t], 2 * perm[t] + 1] for t in range(4)]).ravel()
    assert np.allclose(K0[np.ix_(dof, dof)], KE, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("h", [(1.0, 1.0), (0.5, 0.5), (2.0, 0.5)])
def test_K0_invariants_and_plane_stress_energy(h):
    E, nu = 1.7, 0.3
    lam, mu = fr.plane_stress_lame(E, nu)
    K0 = o2.element_stiffness(E, nu, h)
    assert np.allclose(K0, K0.T, rtol=0, atol=1e-15)
    w = np.linalg.eigvalsh(K0)
    assert np.sum(np.abs(w) < 1e-12 * w.max()) == 3  # two translations and one rotation
    X = np.array([[((l >> 1) & 1) * h[0], (l & 1) * h[1]] for l in range(4)])
    for m in (np.tile([1.0, 0.0], 4), np.tile([0.0, 1.0], 4), np.stack([-X[:, 1], X[:, 0]], 1).ravel()):
        assert np.abs(K0 @ m).max() < 1e-13
    for _ in range(5):
        A = RNG.normal(size=(2, 2))
        u = (X @ A.T).ravel()
        eps = 0.5 * (A + A.T)
        expect = h[0] * h[1] * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
        assert u @ K0 @ u == pytest.approx(expect, rel=1e-13)


# ------------------------------------------------------------------------- K u
@pytest.mark.parametrize("ne", [(2, 2), (3, 5), (6, 4)])
def test_applyK_matches_assembled(ne):
    s = make_sim(ne, np.zeros(2), np.array([2.0, 1.5]))
    K = fr.assemble_K(ne, s.K0(), s.E())
    u = RNG.normal(size=(s.nnodes, 2))
    ref = fr.from_dof(K @ fr.to_dof(u))
    assert np.abs(s.apply_K(u) - ref).max() <= 1e-13 * np.abs(ref).max()
    assert 0.5 * np.sum(u * s.apply_K(u)) == pytest.approx(s.energy_density(u).sum(), rel=1e-12)


# ------------------------------------------------------------------------- multigrid operators
def test_level_operators_are_galerkin_products():
    ne = (16, 8)
    s = make_sim(ne)
    mg = o2.MG(s, 3)
    mg.update_stiffness()
    ops = [fr.assemble_K(ne, s.K0(), s.E())]
    nn = [n + 1 for n in ne]
    for l in range(3):
        nn = [(n - 1) // 2 + 1 for n in nn]
        P = fr.prolongation(nn)
        ops.append((P.T @ ops[-1] @ P).tocsr())
    for l in (1, 2, 3):
        u = RNG.normal(size=(mg.num_nodes(l), 2))
        ref = fr.from_dof(ops[l] @ fr.to_dof(u))
        assert np.abs(mg.apply_K(l, u) - ref).max() <= 1e-12 * np.abs(ref).max(), l
    s1 = o2.Sim((4, 4))
    s1.set_etensor(1.0, 0.3)
    assert np.allclose(o2.MG(s1, 1).chat().sum(0), o2.element_stiffness(1.0, 0.3, (2.0, 2.0)), rtol=1e-13,
                       atol=1e-15)


def test_restriction_is_prolongation_transpose():
    ne = (8, 6)
    s = make_sim(ne)
    mg = o2.MG(s, 1)
    nf, nc = mg.num_nodes(0), mg.num_nodes(1)
    r, xc = RNG.normal(size=(nf, 2)), RNG.normal(size=(nc, 2))
    assert np.sum(mg.restrict(0, r) * xc) == pytest.approx(np.sum(r * mg.interpolate(0, xc)), rel=1e-13)
    P = fr.prolongation([n // 2 + 1 for n in ne])
    assert np.allclose(fr.from_dof(P @ fr.to_dof(xc)), mg.interpolate(0, xc), rtol=0, atol=1e-14)
    base = RNG.normal(size=(nf, 2))
    assert np.allclose(mg.interpolate(0, xc, base, accumulate=True), base + mg.interpolate(0, xc), atol=1e-14)


def test_dirichlet_coarsening_edge_and_point():
    ne = (8, 4)
    s = make_sim(ne)
    nny = ne[1] + 1
    mask = np.zeros(s.nnodes, np.uint8)
    ix, iy = np.arange(s.nnodes) // nny, np.arange(s.nnodes) % nny
    mask[ix == 0] = 3
    mask[(ix == 5) & (iy == 0)] = 2  # odd x index: both coarse neighbours get the y bit
    s.set_dirichlet(mask)
    mg = o2.MG(s, 1)
    m = mg.dirichlet_mask(1)
    cx, cy = np.arange(len(m)) // 3, np.arange(len(m)) % 3
    assert np.all(m[cx == 0] == 3)
    assert m[(cx == 2) & (cy == 0)][0] == 2 and m[(cx == 3) & (cy == 0)][0] == 2
    assert np.count_nonzero(m) == 3 + 2


def test_gauss_seidel_matches_sequential_definition():
    ne = (8, 8)
    s = make_sim(ne)
    nny = ne[1] + 1
    idx = np.arange(s.nnodes)
    ix, iy = idx // nny, idx % nny
    mask = np.zeros(s.nnodes, np.uint8)
    mask[ix == 0] = 3
    mask[(ix == ne[0]) & (iy == 0)] = 2  # roller
    mask[(iy == 4) & (ix == 3)] |= 1
    s.set_dirichlet(mask)
    mg = o2.MG(s, 3)
    mg.update_stiffness()
    ops = [fr.assemble_K(ne, s.K0(), s.E())]
    nn = [n + 1 for n in ne]
    for l in range(2):
        nn = [(n - 1) // 2 + 1 for n in nn]
        P = fr.prolongation(nn)
        ops.append((P.T @ ops[-1] @ P).tocsr())
    for l in (0, 1, 2):
        dm = mg.dirichlet_mask(l)
        nl = len(dm)
        nyl = (ne[1] >> l) + 1
        u = RNG.normal(size=(nl, 2))
        for c in range(2):
            u[(dm >> c) & 1 == 1, c] = 0.0
        b = RNG.normal(size=(nl, 2))
        i_ = np.arange(nl)
        color = 2 * ((i_ // nyl) % 2) + (i_ % nyl) % 2
        for forward in (True, False):
            order = range(4) if forward else range(3, -1, -1)
            ref = fr.from_dof(fr.block_gauss_seidel(ops[l], fr.to_dof(u), fr.to_dof(b),
                                                    [i_[color == c] for c in order], dm, forward))
            out = mg.smooth(l, u, b, forward)
            assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max(), (l, forward)


def _cantilever(ne, levels, rho=None):
    s = make_sim(ne, rho=rho)
    nny = ne[1] + 1
    idx = np.arange(s.nnodes)
    ix, iy = idx // nny, idx % nny
    mask = np.where(ix == 0, 3, 0).astype(np.uint8)
    s.set_dirichlet(mask)
    f = np.zeros((s.nnodes, 2))
    f[(ix == ne[0]) & (iy == 0), 1] = -1.0
    s.set_forces(f)
    return s, o2.MG(s, levels), mask, f


def _reduced_solve(K, mask, f):
    fixed = np.zeros(K.shape[0], bool)
    for c in range(2):
        fixed[c::2] |= ((mask >> c) & 1).astype(bool)
    free = ~fixed
    u = np.zeros(K.shape[0])
    u[free] = spla.spsolve(K[free][:, free].tocsc(), fr.to_dof(f)[free])
    return u


def test_coarse_solve_matches_sparse_direct():
    ne = (16, 8)
    s, mg, mask, f = _cantilever(ne, 2)
    mg.update_stiffness()
    b = RNG.normal(size=(mg.num_nodes(2), 2))
    K = fr.assemble_K(ne, s.K0(), s.E())
    for nn in ([9, 5], [5, 3]):
        P = fr.prolongation(nn)
        K = (P.T @ K @ P).tocsr()
    ref = _reduced_solve(K, mg.dirichlet_mask(2), b)
    assert np.abs(fr.to_dof(mg.coarse_solve(b)) - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("levels,fmg,nu", [(1, False, 1), (2, False, 1), (3, True, 2), (4, False, 2)])
def test_pcg_converges_to_direct_solution(levels, fmg, nu):
    ne = (32, 16)
    s, mg, mask, f = _cantilever(ne, levels)
    u, its = mg.pcg(np.zeros((s.nnodes, 2)), f, 300, 1e-12, 1, nu, fmg)
    ref = _reduced_solve(fr.assemble_K(ne, s.K0(), s.E()), mask, f)
    assert np.abs(fr.to_dof(u) - ref).max() <= 1e-9 * np.abs(ref).max()
    assert its < 80


def test_odd_grid_and_nan():
    with pytest.raises(RuntimeError, match="divisible by 2"):
        o2.MG(make_sim((6, 4)), 2)
    s, mg, mask, f = _cantilever((8, 4), 1)
    bad = f.copy()
    bad[-3, 0] = np.nan
    with pytest.raises(RuntimeError, match="NaN encountered"):
        mg.pcg(np.zeros((s.nnodes, 2)), bad, 10, 1e-8)


# ------------------------------------------------------------------------- sensitivities
@pytest.mark.parametrize("law,gravity", [(0, False), (1, False), (0, True), (1, True)])
def test_compliance_gradient_matches_central_fd(law, gravity):
    ne = (4, 3)
    rho = RNG.uniform(0.3, 0.9, 12)
    s = make_sim(ne, rho=rho)
    s.set_law(law, 1.0, 1e-9, 3.0, 3.0)
    nny = ne[1] + 1
    idx = np.arange(s.nnodes)
    ix, iy = idx // nny, idx % nny
    if gravity:
        mask = np.where(iy == 0, 3, 0).astype(np.uint8)
        s.set_gravity([0.3, -1.0])
    else:
        mask = np.where(ix == 0, 3, 0).astype(np.uint8)
        f = np.zeros((s.nnodes, 2))
        f[(ix == ne[0]) & (iy == 0), 1] = -1.0
        s.set_forces(f)
    s.set_dirichlet(mask)

    def comp(r):
        s.set_densities(r)
        f = s.build_load()
        u = _reduced_solve(fr.assemble_K(ne, s.K0(), s.E()), mask, f)
        return 0.5 * fr.to_dof(f) @ u, u

    _, u = comp(rho)
    if gravity:
        assert s.build_load()[:, 1].sum() == pytest.approx(-rho.sum(), rel=1e-14)  # unit element areas
    g = s.compliance_gradient(fr.from_dof(u))
    h = 1e-6
    for e in range(12):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        fd = (comp(rp)[0] - comp(rm)[0]) / (2 * h)
        assert g[e] == pytest.approx(fd, rel=2e-6, abs=1e-10), e


# ------------------------------------------------------------------------- filter (TopologyOptimizationFilter.hh:297-400)
def _dense_filter2d(dims, rmin, linear):
    n = dims[0] * dims[1]
    R = int(np.ceil(rmin)) - 1
    A = np.zeros((n, n))

    def refl(i, s):
        while i < 0 or i >= s:
            i = 2 * s - i - 1 if i >= s else i
            i = -i - 1 if i < 0 else i
        return i

    for i in range(dims[0]):
        for j in range(dims[1]):
            for a in range(-R, R + 1):
                for b in range(-R, R + 1):
                    w = rmin - np.hypot(a, b) if linear else 1.0
                    if w > 0:
                        A[i * dims[1] + j, refl(i + a, dims[0]) * dims[1] + refl(j + b, dims[1])] += w
    return A / A.sum(1, keepdims=True)


@pytest.mark.parametrize("rmin,linear", [(2.0, True), (4.0, True), (3.0, False)])
def test_smoothing_filter_2d(rmin, linear):
    dims = (7, 5)
    x = RNG.uniform(size=35)
    A = _dense_filter2d(dims, rmin, linear)
    assert np.allclose(o2.smoothing_filter(dims, rmin, int(linear), x), A @ x, rtol=1e-13, atol=1e-14)
    # a 2D Linear filter of radius 1 weighs the 4 edge neighbours 1 and the diagonals 2 - sqrt(2)
    x1 = np.zeros(35)
    x1[2 * 5 + 2] = 1.0
    out = o2.smoothing_filter(dims, 2.0, 1, x1)
    W = 2 + 4 + 4 * (2 - np.sqrt(2))
    assert out[2 * 5 + 2] == pytest.approx(2 / W, rel=1e-14)
    assert out[2 * 5 + 3] == pytest.approx(1 / W, rel=1e-14)
    assert out[3 * 5 + 3] == pytest.approx((2 - np.sqrt(2)) / W, rel=1e-14)


# ------------------------------------------------------------------------- layer-by-layer (LayerByLayer.hh:223-300)
def test_masked_layer_solves_equal_truncated_grids():
    ne = (8, 6)
    rho = RNG.uniform(0.3, 1.0, 48)
    s = make_sim(ne, rho=rho)
    nny = ne[1] + 1
    iy = np.arange(s.nnodes) % nny
    s.set_dirichlet(np.where(iy == 0, 3, 0).astype(np.uint8))
    s.set_gravity([0.0, -1.0])
    mg = o2.MG(s, 1)
    obj, grad, comp, its = mg.lbl_run_zero(2, 200, 1e-11)
    comps = []
    for k, l in enumerate([6, 4, 2]):
        t = make_sim((8, l), rho=rho.reshape(ne)[:, :l].ravel())
        tmask = np.where(np.arange(t.nnodes) % (l + 1) == 0, 3, 0).astype(np.uint8)
        t.set_dirichlet(tmask)
        t.set_gravity([0.0, -1.0])
        f = t.build_load()
        u = _reduced_solve(fr.assemble_K((8, l), t.K0(), t.E()), tmask, f)
        comps.append(fr.to_dof(f) @ u)
        assert comp[k] == pytest.approx(comps[-1], rel=1e-9)
    assert obj == pytest.approx(0.5 * sum(comps) / 3, rel=1e-9)
