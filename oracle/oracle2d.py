"""ctypes front end of the planar (2D) CPU oracle (oracle/vf_oracle2d.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the product. Nodal fields are numpy
arrays of shape ``(nn, 2)`` (the reference's ColMajor VField at N = 2), stored as the two component planes
the C code expects.
"""
import ctypes

import numpy as np

from .oracle import _d, _l, _p, _u8, lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc2_last_error().decode())


def to_planes(u):
    return np.ascontiguousarray(np.asarray(u, dtype=np.float64).T).reshape(-1)


def from_planes(v, nn):
    return np.ascontiguousarray(v.reshape(2, nn).T)


def _setup():
    L = lib()
    if getattr(L, "_orc2_ready", False):
        return L
    L.orc2_last_error.restype = ctypes.c_char_p
    L.orc2_sim_create.restype = ctypes.c_void_p
    L.orc2_sim_create.argtypes = [ctypes.c_long, ctypes.c_long, _d, _d]
    L.orc2_mg_create.restype = ctypes.c_void_p
    L.orc2_mg_create.argtypes = [ctypes.c_void_p, ctypes.c_long]
    L._orc2_ready = True
    return L


def element_stiffness(E, nu, h):
    out = np.zeros(64)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    _setup().orc2_element_stiffness(ctypes.c_double(E), ctypes.c_double(nu), _p(hh), _p(out))
    return out.reshape(8, 8)


def smoothing_filter(dims, rmin, type_, x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    out = np.zeros_like(x)
    d = np.ascontiguousarray(dims, dtype=np.int64)
    _check(_setup().orc2_smoothing_filter(d.ctypes.data_as(_l), ctypes.c_double(rmin), ctypes.c_int(type_), _p(x),
                                          _p(out)))
    return out


class Sim:
    """Planar simulator state (TensorProductSimulator<double,1,1> restated)."""

    def __init__(self, ne, lo=None, hi=None):
        self.ne = tuple(int(x) for x in ne)
        lo = np.zeros(2) if lo is None else np.asarray(lo, dtype=np.float64)
        hi = np.asarray(self.ne, dtype=np.float64) if hi is None else np.asarray(hi, dtype=np.float64)
        self.lo, self.hi = lo.copy(), hi.copy()
        self.h = (hi - lo) / np.asarray(self.ne)
        self.nn = tuple(x + 1 for x in self.ne)
        self.nnodes = int(np.prod(self.nn))
        self.nelems = int(np.prod(self.ne))
        self._h = _setup().orc2_sim_create(*self.ne, _p(np.ascontiguousarray(lo)), _p(np.ascontiguousarray(hi)))
        if not self._h:
            raise RuntimeError(lib().orc2_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc2_sim_destroy(ctypes.c_void_p(self._h))
            self._h = None

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def set_etensor(self, E, nu):
        _check(lib().orc2_sim_set_etensor(self.handle, ctypes.c_double(E), ctypes.c_double(nu)))

    def set_law(self, law=0, E0=1.0, Emin=1e-4, gamma=3.0, q=3.0):
        _check(lib().orc2_sim_set_law(self.handle, ctypes.c_int(law), ctypes.c_double(E0), ctypes.c_double(Emin),
                                      ctypes.c_double(gamma), ctypes.c_double(q)))

    def set_densities(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        assert rho.size == self.nelems
        _check(lib().orc2_sim_set_densities(self.handle, _p(rho)))

    def E(self):
        out = np.zeros(self.nelems)
        _check(lib().orc2_sim_get_E(self.handle, _p(out)))
        return out

    def set_E(self, E):
        """Replace the moduli (after set_law / set_densities, which recompute them)."""
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1)
        assert E.size == self.nelems
        _check(lib().orc2_sim_set_E(self.handle, _p(E)))

    def K0(self):
        out = np.zeros(64)
        _check(lib().orc2_sim_get_K0(self.handle, _p(out)))
        return out.reshape(8, 8)

    def set_dirichlet(self, mask, values=None):
        """mask: (nn,) uint8 component bits or (nn, 2) bool."""
        mask = np.asarray(mask)
        if mask.ndim == 2:
            mask = mask[:, 0].astype(np.uint8) | (mask[:, 1].astype(np.uint8) << 1)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        vals = None if values is None else to_planes(values)
        _check(lib().orc2_sim_set_dirichlet(self.handle, mask.ctypes.data_as(_u8), None if vals is None else _p(vals)))

    def set_forces(self, f):
        _check(lib().orc2_sim_set_forces(self.handle, _p(to_planes(f))))

    def set_gravity(self, g):
        _check(lib().orc2_sim_set_gravity(self.handle, _p(np.ascontiguousarray(g, dtype=np.float64))))

    def set_mask_height(self, h):
        _check(lib().orc2_sim_set_mask_height(self.handle, ctypes.c_double(h)))

    def apply_K(self, u):
        up = to_planes(u)
        y = np.zeros_like(up)
        _check(lib().orc2_sim_apply_K(self.handle, _p(up), _p(y)))
        return from_planes(y, self.nnodes)

    def build_load(self):
        f = np.zeros(2 * self.nnodes)
        _check(lib().orc2_sim_build_load(self.handle, _p(f)))
        return from_planes(f, self.nnodes)

    def compliance_gradient(self, u):
        g = np.zeros(self.nelems)
        _check(lib().orc2_sim_compliance_gradient(self.handle, _p(to_planes(u)), _p(g)))
        return g

    def energy_density(self, u):
        g = np.zeros(self.nelems)
        _check(lib().orc2_sim_energy_density(self.handle, _p(to_planes(u)), _p(g)))
        return g

    def masked_dot(self, a, b):
        out = ctypes.c_double()
        _check(lib().orc2_sim_masked_dot(self.handle, _p(to_planes(a)), _p(to_planes(b)), ctypes.byref(out)))
        return out.value


class MG:
    """MultigridSolver<double,1,1> restated; holds a reference to its fine Sim."""

    def __init__(self, sim, levels):
        self.sim = sim
        self.levels = int(levels)
        self._h = _setup().orc2_mg_create(sim.handle, ctypes.c_long(self.levels))
        if not self._h:
            raise RuntimeError(lib().orc2_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc2_mg_destroy(ctypes.c_void_p(self._h))
            self._h = None

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def num_nodes(self, l):
        out = ctypes.c_long()
        _check(lib().orc2_mg_num_nodes(self.handle, ctypes.c_long(l), ctypes.byref(out)))
        return out.value

    def dirichlet_mask(self, l):
        out = np.zeros(self.num_nodes(l), dtype=np.uint8)
        _check(lib().orc2_mg_dirichlet_mask(self.handle, ctypes.c_long(l), out.ctypes.data_as(_u8)))
        return out

    def set_symmetric(self, s):
        _check(lib().orc2_mg_set_symmetric(self.handle, ctypes.c_int(int(bool(s)))))

    def set_mask_layer(self, layer):
        _check(lib().orc2_mg_set_mask_layer(self.handle, ctypes.c_long(int(layer))))

    def update_stiffness(self):
        _check(lib().orc2_mg_update_stiffness(self.handle))

    def apply_K(self, l, u):
        nn = self.num_nodes(l)
        y = np.zeros(2 * nn)
        _check(lib().orc2_mg_apply_K(self.handle, ctypes.c_long(l), _p(to_planes(u)), _p(y)))
        return from_planes(y, nn)

    def residual(self, l, u, b):
        nn = self.num_nodes(l)
        r = np.zeros(2 * nn)
        _check(lib().orc2_mg_residual(self.handle, ctypes.c_long(l), _p(to_planes(u)), _p(to_planes(b)), _p(r)))
        return from_planes(r, nn)

    def smooth(self, l, u, b, forward=True):
        nn = self.num_nodes(l)
        up = to_planes(u)
        _check(lib().orc2_mg_smooth(self.handle, ctypes.c_long(l), _p(up), _p(to_planes(b)), ctypes.c_int(int(forward))))
        return from_planes(up, nn)

    def restrict(self, l, rf):
        nc = self.num_nodes(l + 1)
        rc = np.zeros(2 * nc)
        _check(lib().orc2_mg_restrict(self.handle, ctypes.c_long(l), _p(to_planes(rf)), _p(rc)))
        return from_planes(rc, nc)

    def interpolate(self, l, xc, xf=None, accumulate=False):
        nf = self.num_nodes(l)
        xfp = np.zeros(2 * nf) if xf is None else to_planes(xf)
        _check(lib().orc2_mg_interpolate(self.handle, ctypes.c_long(l), _p(to_planes(xc)), _p(xfp),
                                         ctypes.c_int(int(accumulate))))
        return from_planes(xfp, nf)

    def coarse_solve(self, b):
        nc = self.num_nodes(self.levels)
        x = np.zeros(2 * nc)
        _check(lib().orc2_mg_coarse_solve(self.handle, _p(to_planes(b)), _p(x)))
        return from_planes(x, nc)

    def blocks(self, l):
        out = np.zeros(self.num_nodes(l) * 9 * 4)
        _check(lib().orc2_mg_blocks(self.handle, ctypes.c_long(l), _p(out)))
        return out.reshape(-1, 9, 2, 2)

    def element_matrices(self, l):
        ne = self.sim.nelems // (4 ** l)
        out = np.zeros(ne * 64)
        _check(lib().orc2_mg_element_matrices(self.handle, ctypes.c_long(l), _p(out)))
        return out.reshape(ne, 8, 8)

    def chat(self):
        out = np.zeros(4 * 64)
        _check(lib().orc2_mg_chat(self.handle, _p(out)))
        return out.reshape(4, 8, 8)

    def precondition(self, r, mg_iter=1, nu=1, fmg=False):
        nn = self.num_nodes(0)
        s = np.zeros(2 * nn)
        _check(lib().orc2_mg_precondition(self.handle, _p(to_planes(r)), _p(s), ctypes.c_long(mg_iter),
                                          ctypes.c_long(nu), ctypes.c_int(int(fmg))))
        return from_planes(s, nn)

    def pcg(self, x, b, max_iter, tol, mg_iter=1, nu=1, fmg=False, dirichlet_satisfied=False, hist_len=0):
        nn = self.num_nodes(0)
        xp, bp = to_planes(x), to_planes(b)
        hist = np.zeros(max(hist_len, 1))
        its = ctypes.c_long()
        _check(lib().orc2_mg_pcg(self.handle, _p(xp), _p(bp), ctypes.c_long(max_iter), ctypes.c_double(tol),
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
        nl = (s.ne[1] + layer_increment - 1) // layer_increment
        comp = np.zeros(nl)
        its = np.zeros(nl, dtype=np.int64)
        obj = ctypes.c_double()
        _check(lib().orc2_lbl_run_zero(self.handle, ctypes.c_long(layer_increment), ctypes.c_long(max_iter),
                                       ctypes.c_double(tol), ctypes.c_long(mg_iter), ctypes.c_long(nu),
                                       ctypes.c_int(int(fmg)), ctypes.byref(obj), _p(grad), _p(comp),
                                       its.ctypes.data_as(_l)))
        return obj.value, grad, comp, its
