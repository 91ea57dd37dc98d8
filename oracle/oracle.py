"""ctypes front end of the CPU oracle (oracle/vf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of bench.py, always as the checker, never as the thing measured or
shipped. Field layout follows the reference's VField (ColMajor ``(numNodes, 3)``): this
wrapper takes and returns numpy arrays of shape ``(nn, 3)`` and stores them as the
component planes the C code expects.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_l = ctypes.POINTER(ctypes.c_long)


def build(force=False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])


def use_library(path):
    """Load the oracle from `path` (a copy built elsewhere, e.g. tuned for the timing host) instead of the
    in-tree library; must be called before the first use."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("oracle library already loaded")
    _LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_sim_create.restype = ctypes.c_void_p
        L.orc_sim_create.argtypes = [ctypes.c_long] * 3 + [_d, _d]
        L.orc_mg_create.restype = ctypes.c_void_p
        L.orc_mg_create.argtypes = [ctypes.c_void_p, ctypes.c_long]
        for name in dir(L):
            pass
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_d)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def to_planes(u):
    u = np.asarray(u, dtype=np.float64)
    return np.ascontiguousarray(u.T).reshape(-1)


def from_planes(v, nn):
    return np.ascontiguousarray(v.reshape(3, nn).T)


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(int(n)))


def get_threads():
    return int(lib().orc_get_threads())


def element_stiffness(E, nu, h):
    out = np.zeros(576)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    lib().orc_element_stiffness(ctypes.c_double(E), ctypes.c_double(nu), _p(hh), _p(out))
    return out.reshape(24, 24)


class Sim:
    """Fine-grid simulator state (TensorProductSimulator restated)."""

    def __init__(self, ne, lo=None, hi=None):
        self.ne = tuple(int(x) for x in ne)
        lo = np.zeros(3) if lo is None else np.asarray(lo, dtype=np.float64)
        hi = np.asarray(self.ne, dtype=np.float64) if hi is None else np.asarray(hi, dtype=np.float64)
        self.lo, self.hi = lo.copy(), hi.copy()
        self.h = (hi - lo) / np.asarray(self.ne)
        self.nn = tuple(x + 1 for x in self.ne)
        self.nnodes = int(np.prod(self.nn))
        self.nelems = int(np.prod(self.ne))
        self._h = lib().orc_sim_create(*self.ne, _p(np.ascontiguousarray(lo)), _p(np.ascontiguousarray(hi)))
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_sim_destroy(ctypes.c_void_p(self._h))
            self._h = None

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def set_etensor(self, E, nu):
        _check(lib().orc_sim_set_etensor(self.handle, ctypes.c_double(E), ctypes.c_double(nu)))

    def set_law(self, law=0, E0=1.0, Emin=1e-4, gamma=3.0, q=3.0):
        _check(lib().orc_sim_set_law(self.handle, ctypes.c_int(law), ctypes.c_double(E0), ctypes.c_double(Emin),
                                     ctypes.c_double(gamma), ctypes.c_double(q)))

    def set_densities(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        assert rho.size == self.nelems
        _check(lib().orc_sim_set_densities(self.handle, _p(rho)))

    def E(self):
        out = np.zeros(self.nelems)
        _check(lib().orc_sim_get_E(self.handle, _p(out)))
        return out

    def K0(self):
        out = np.zeros(576)
        _check(lib().orc_sim_get_K0(self.handle, _p(out)))
        return out.reshape(24, 24)

    def set_dirichlet(self, mask, values=None):
        """mask: (nn,) uint8 component bits or (nn,3) bool."""
        mask = np.asarray(mask)
        if mask.ndim == 2:
            mask = (mask[:, 0].astype(np.uint8) | (mask[:, 1].astype(np.uint8) << 1) | (mask[:, 2].astype(np.uint8) << 2))
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        vals = None if values is None else to_planes(values)
        _check(lib().orc_sim_set_dirichlet(self.handle, mask.ctypes.data_as(_u8), None if vals is None else _p(vals)))

    def set_forces(self, f):
        fp = to_planes(f)
        _check(lib().orc_sim_set_forces(self.handle, _p(fp)))

    def set_gravity(self, g):
        gg = np.ascontiguousarray(g, dtype=np.float64)
        _check(lib().orc_sim_set_gravity(self.handle, _p(gg)))

    def apply_K(self, u):
        up = to_planes(u)
        y = np.zeros_like(up)
        _check(lib().orc_sim_apply_K(self.handle, _p(up), _p(y)))
        return from_planes(y, self.nnodes)

    def build_load(self):
        f = np.zeros(3 * self.nnodes)
        _check(lib().orc_sim_build_load(self.handle, _p(f)))
        return from_planes(f, self.nnodes)

    def compliance_gradient(self, u):
        up = to_planes(u)
        g = np.zeros(self.nelems)
        _check(lib().orc_sim_compliance_gradient(self.handle, _p(up), _p(g)))
        return g

    def energy_density(self, u):
        up = to_planes(u)
        g = np.zeros(self.nelems)
        _check(lib().orc_sim_energy_density(self.handle, _p(up), _p(g)))
        return g


class MG:
    """MultigridSolver restated; holds a reference to its fine Sim."""

    def __init__(self, sim, levels):
        self.sim = sim
        self.levels = int(levels)
        self._h = lib().orc_mg_create(sim.handle, ctypes.c_long(self.levels))
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mg_destroy(ctypes.c_void_p(self._h))
            self._h = None

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def num_nodes(self, l):
        out = ctypes.c_long()
        _check(lib().orc_mg_num_nodes(self.handle, ctypes.c_long(l), ctypes.byref(out)))
        return out.value

    def dirichlet_mask(self, l):
        out = np.zeros(self.num_nodes(l), dtype=np.uint8)
        _check(lib().orc_mg_dirichlet_mask(self.handle, ctypes.c_long(l), out.ctypes.data_as(_u8)))
        return out

    def set_symmetric(self, s):
        _check(lib().orc_mg_set_symmetric(self.handle, ctypes.c_int(int(bool(s)))))

    def set_mask_layer(self, layer):
        _check(lib().orc_mg_set_mask_layer(self.handle, ctypes.c_long(int(layer))))

    def update_stiffness(self):
        _check(lib().orc_mg_update_stiffness(self.handle))

    def apply_K(self, l, u):
        nn = self.num_nodes(l)
        up = to_planes(u)
        y = np.zeros(3 * nn)
        _check(lib().orc_mg_apply_K(self.handle, ctypes.c_long(l), _p(up), _p(y)))
        return from_planes(y, nn)

    def residual(self, l, u, b):
        nn = self.num_nodes(l)
        up, bp = to_planes(u), to_planes(b)
        r = np.zeros(3 * nn)
        _check(lib().orc_mg_residual(self.handle, ctypes.c_long(l), _p(up), _p(bp), _p(r)))
        return from_planes(r, nn)

    def smooth(self, l, u, b, forward=True):
        nn = self.num_nodes(l)
        up, bp = to_planes(u), to_planes(b)
        _check(lib().orc_mg_smooth(self.handle, ctypes.c_long(l), _p(up), _p(bp), ctypes.c_int(int(forward))))
        return from_planes(up, nn)

    def restrict(self, l, rf):
        nc = self.num_nodes(l + 1)
        rp = to_planes(rf)
        rc = np.zeros(3 * nc)
        _check(lib().orc_mg_restrict(self.handle, ctypes.c_long(l), _p(rp), _p(rc)))
        return from_planes(rc, nc)

    def interpolate(self, l, xc, xf=None, accumulate=False):
        nf = self.num_nodes(l)
        xcp = to_planes(xc)
        xfp = np.zeros(3 * nf) if xf is None else to_planes(xf)
        _check(lib().orc_mg_interpolate(self.handle, ctypes.c_long(l), _p(xcp), _p(xfp), ctypes.c_int(int(accumulate))))
        return from_planes(xfp, nf)

    def coarse_solve(self, b):
        nc = self.num_nodes(self.levels)
        bp = to_planes(b)
        x = np.zeros(3 * nc)
        _check(lib().orc_mg_coarse_solve(self.handle, _p(bp), _p(x)))
        return from_planes(x, nc)

    def blocks(self, l):
        out = np.zeros(self.num_nodes(l) * 27 * 9)
        _check(lib().orc_mg_blocks(self.handle, ctypes.c_long(l), _p(out)))
        return out.reshape(-1, 27, 3, 3)

    def chat(self):
        out = np.zeros(8 * 576)
        _check(lib().orc_mg_chat(self.handle, _p(out)))
        return out.reshape(8, 24, 24)

    def precondition(self, r, mg_iter=1, nu=1, fmg=False):
        nn = self.num_nodes(0)
        rp = to_planes(r)
        s = np.zeros(3 * nn)
        _check(lib().orc_mg_precondition(self.handle, _p(rp), _p(s), ctypes.c_long(mg_iter), ctypes.c_long(nu),
                                         ctypes.c_int(int(fmg))))
        return from_planes(s, nn)

    def pcg(self, x, b, max_iter, tol, mg_iter=1, nu=1, fmg=False, dirichlet_satisfied=False, hist_len=0):
        nn = self.num_nodes(0)
        xp, bp = to_planes(x), to_planes(b)
        hist = np.zeros(max(hist_len, 1))
        its = ctypes.c_long()
        _check(lib().orc_mg_pcg(self.handle, _p(xp), _p(bp), ctypes.c_long(max_iter), ctypes.c_double(tol),
                                ctypes.c_long(mg_iter), ctypes.c_long(nu), ctypes.c_int(int(fmg)),
                                ctypes.c_int(int(dirichlet_satisfied)), _p(hist), ctypes.c_long(hist_len),
                                ctypes.byref(its)))
        out = from_planes(xp, nn)
        if hist_len:
            return out, its.value, hist[: min(its.value, hist_len)]
        return out, its.value

    def lbl_run_zero(self, layer_increment, max_iter, tol, mg_iter=1, nu=1, fmg=False):
        s = self.sim
        grad = np.zeros(s.nelems)
        nl = (s.ne[1] + layer_increment - 1) // layer_increment  # layer solves: heights ne1, ne1 - inc, ...
        comp = np.zeros(nl)
        its = np.zeros(nl, dtype=np.int64)
        obj = ctypes.c_double()
        _check(lib().orc_lbl_run_zero(self.handle, ctypes.c_long(layer_increment), ctypes.c_long(max_iter),
                                      ctypes.c_double(tol), ctypes.c_long(mg_iter), ctypes.c_long(nu),
                                      ctypes.c_int(int(fmg)), ctypes.byref(obj), _p(grad), _p(comp),
                                      its.ctypes.data_as(_l)))
        return obj.value, grad, comp, its


def smoothing_filter(x, dims, rmin, linear=True):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    out = np.zeros_like(x)
    d = np.ascontiguousarray(dims, dtype=np.int64)
    _check(lib().orc_smoothing_filter(d.ctypes.data_as(_l), ctypes.c_double(rmin), ctypes.c_int(int(linear)),
                                      _p(x), _p(out)))
    return out


def projection(x, beta):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    out = np.zeros_like(x)
    _check(lib().orc_projection(ctypes.c_double(beta), ctypes.c_long(x.size), _p(x), _p(out)))
    return out


def projection_backprop(g, vars_, beta):
    g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
    v = np.ascontiguousarray(vars_, dtype=np.float64).reshape(-1)
    out = np.zeros_like(g)
    _check(lib().orc_projection_backprop(ctypes.c_double(beta), ctypes.c_long(g.size), _p(g), _p(v), _p(out)))
    return out


def oc_update(x0, dJ, dc, lam, m=0.2, p=0.5):
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    dJ = np.ascontiguousarray(dJ, dtype=np.float64)
    dc = np.ascontiguousarray(dc, dtype=np.float64)
    out = np.zeros_like(x0)
    _check(lib().orc_oc_update(ctypes.c_long(x0.size), _p(x0), _p(dJ), _p(dc), ctypes.c_double(lam),
                               ctypes.c_double(m), ctypes.c_double(p), _p(out)))
    return out


# ------------------------------------------------------------------------------------------------
# Filter chain, volume constraint and one optimality-criterion step, composed from the primitives
# above (TopologyOptimizationFilter.hh:142-175, TopologyOptimizationConstraint.hh:30-37,
# OptimalityCriterion.hh:51-138).
class Chain:
    def __init__(self, dims, rmin=0.0, linear=True, beta=0.0):
        self.dims, self.rmin, self.linear, self.beta = tuple(dims), rmin, linear, beta

    def forward(self, x):
        s = smoothing_filter(x, self.dims, self.rmin, self.linear) if self.rmin > 0 else np.asarray(x, float).copy()
        p = projection(s, self.beta) if self.beta > 0 else s
        return s, p

    def backprop(self, g, x):
        s, _ = self.forward(x)
        if self.beta > 0:
            g = projection_backprop(g, s, self.beta)
        if self.rmin > 0:
            g = smoothing_filter(g, self.dims, self.rmin, self.linear)
        return g


class OC:
    """OCOptimizer restated: the lambda bracket persists across steps (OptimalityCriterion.hh:47-48)."""

    def __init__(self):
        self.lmin, self.lmax = 1.0, 2.0

    def step(self, x0, dJ, dc, vf, chain, m=0.2, p=0.5, ctol=1e-6, guard=100):
        def ceval(lam):
            xs = oc_update(x0, dJ, dc, lam, m, p)
            return 1.0 - chain.forward(xs)[1].mean() / vf, xs

        dil = 32.0
        mid = 0.5 * (self.lmin + self.lmax)
        self.lmax = dil * self.lmax + (1 - dil) * mid
        self.lmin = max(dil * self.lmin + (1 - dil) * mid, 0.01)
        nit = 0
        while nit < guard:
            if ceval(self.lmin)[0] < 0:
                break
            self.lmax = self.lmin
            self.lmin /= 2
            nit += 1
        if nit == guard:
            raise RuntimeError("Bracketing constraint(lambda_min) < 0 failed")
        if nit == 0:
            while nit < guard:
                if ceval(self.lmax)[0] > 0:
                    break
                self.lmin = self.lmax
                self.lmax *= 2
                nit += 1
        if nit == guard:
            raise RuntimeError("Bracketing constraint(lambda_max) > 0 failed")
        while True:
            mid = 0.5 * (self.lmin + self.lmax)
            v, xs = ceval(mid)
            if abs(v) <= ctol:
                return xs
            if v < 0:
                self.lmin = mid
            if v > 0:
                self.lmax = mid


# ---------------------------------------------------------------------------------------------------------
# Boundary utilities of the simulator, restated in numpy (small grids only).

def _node_grid(ne):
    nn = [n + 1 for n in ne]
    ix, iy, iz = np.meshgrid(np.arange(nn[0]), np.arange(nn[1]), np.arange(nn[2]), indexing="ij")
    return nn, ix.ravel(), iy.ravel(), iz.ravel()


def _elem_nodes(ne):
    nn = [n + 1 for n in ne]
    ex, ey, ez = np.meshgrid(np.arange(ne[0]), np.arange(ne[1]), np.arange(ne[2]), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    cols = []
    for l in range(8):
        i, j, k = (l >> 2) & 1, (l >> 1) & 1, l & 1
        cols.append(((ex + i) * nn[1] + ey + j) * nn[2] + ez + k)
    return np.stack(cols, axis=1)


def constant_strain_load(ne, h, young, nu, rho, eps):
    """TensorProductSimulator::constantStrainLoad (TensorProductSimulator.hh:1119-1145) with the element load of
    constantStrainLoad / constantStressLoad (84-113): sigma = C : eps for the isotropic material tensor, element
    e adds rho_e * V * int eps(phi_l e_c) : sigma to component c of its node l, i.e.
    rho_e * sum_d sigma[c][d] * s_d(l) * V / (4 h_d). Returns the (nn, 3) load."""
    eps = np.asarray(eps, dtype=float)
    lam = young * nu / ((1 + nu) * (1 - 2 * nu))
    mu = young / (2 * (1 + nu))
    sigma = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    V = h[0] * h[1] * h[2]
    en = _elem_nodes(ne)
    nn = int(np.prod([n + 1 for n in ne]))
    F = np.zeros((nn, 3))
    rho = np.asarray(rho, dtype=float)
    for l in range(8):
        s = np.array([1.0 if (l >> (2 - d)) & 1 else -1.0 for d in range(3)])
        load_l = sigma @ (s * V / (4.0 * np.asarray(h, dtype=float)))  # (3,)
        np.add.at(F, en[:, l], rho[:, None] * load_l[None, :])
    return F


def sample_nodal_field(ne, lo, hi, u, p):
    """TensorProductSimulator::sampleNodalField (TensorProductSimulator.hh:1161-1179): clamp each point to the
    domain, find its element (getElementNDIndex, 1731-1752) and evaluate the element's trilinear interpolant at
    the clamped point's own coordinates q (the reference passes q, not the reference coordinates of q):
    sum_l u[node(e, l)] prod_d (bit_d(l) ? q_d : 1 - q_d)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    u = np.asarray(u, float)
    if u.ndim == 1:
        u = u[:, None]
    nn = [n + 1 for n in ne]
    h = (hi - lo) / np.asarray(ne, float)
    out = np.zeros((len(p), u.shape[1]))
    for i, pt in enumerate(np.asarray(p, float)):
        q = np.minimum(np.maximum(pt, lo), hi)
        e = []
        for d in range(3):
            f = (q[d] - lo[d]) / h[d]
            e.append(ne[d] - 1 if abs(f - ne[d]) < 1e-10 else int(f))
        val = 0.0
        for l in range(8):
            a, b, c = (l >> 2) & 1, (l >> 1) & 1, l & 1
            w = (q[2] if c else 1 - q[2]) * (q[1] if b else 1 - q[1]) * (q[0] if a else 1 - q[0])
            n = ((e[0] + a) * nn[1] + e[1] + b) * nn[2] + e[2] + c
            val = val + u[n] * w
        out[i] = val
    return out


def stiffness_upper(ne, K0, E):
    """TensorProductSimulator::getK (TensorProductSimulator.hh:1506-1513, 833-865): the upper triangle (row <=
    column, DoF 3 node + component) of sum_e E_e K0, as a scipy CSC matrix."""
    import scipy.sparse as sp

    en = _elem_nodes(ne)
    dofs = (3 * en[:, :, None] + np.arange(3)[None, None, :]).reshape(len(en), 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = (np.asarray(E)[:, None, None] * np.asarray(K0)[None]).reshape(-1)
    keep = rows <= cols
    nn = int(np.prod([n + 1 for n in ne]))
    return sp.csc_matrix((vals[keep], (rows[keep], cols[keep])), shape=(3 * nn, 3 * nn))


def mesh(ne, lo, hi):
    """TensorProductSimulator::getMesh (TensorProductSimulator.hh:747-776): node positions and the hexahedra with
    the local node pairs (0,1), (3,2), (4,5), (7,6) (Gmsh order)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    nn, ix, iy, iz = _node_grid(ne)
    h = (hi - lo) / np.asarray(ne, float)
    V = np.stack([lo[0] + ix * h[0], lo[1] + iy * h[1], lo[2] + iz * h[2]], axis=1)
    en = _elem_nodes(ne)
    F = en[:, [0, 1, 3, 2, 4, 5, 7, 6]]
    return V, F
