"""UpsampleFilter, VertexToCellFilter and LangelaarFilter of TopologyOptimizationFilter.hh, restated in numpy.

TEST INFRASTRUCTURE ONLY: the checker of the product's filter kernels, never shipped or measured.

Grids are C order (last index fastest), 2 or 3 axes; a planar LangelaarFilter grid is (nx, ny) with y the build
direction. Pins (tests/test_filters_oracle.py): upsample against scipy's RegularGridInterpolator and exact
reproduction of multilinear fields; vertex-to-cell against scipy.signal.convolve; every adjoint against the
transpose of its forward map (dense matrices on small grids); Langelaar against the exact min/max additive-
manufacturing filter it smooths (within the smoothing's error) and its backprop against central differences.
"""
import itertools

import numpy as np

# LangelaarFilter constants (TopologyOptimizationFilter.hh:701-710)
LG_EPS, LG_P, LG_Q = 1e-4, 40.0, 40.0 - 1.58


def upsample_dims(in_dims, factor):
    """m_setInputDimensions (TopologyOptimizationFilter.hh:441-445)."""
    return tuple((n - 1) * factor + 1 for n in in_dims)


def upsample(x, factor):
    """UpsampleFilter::apply, degree 1 (TopologyOptimizationFilter.hh:460-488): every coarse cell writes the
    multilinear interpolant of its 2^N corners at the fine nodes it covers (fine node j sits at local coordinate
    (j - f c) / f of cell c)."""
    x = np.asarray(x, float)
    N = x.ndim
    out = np.empty(upsample_dims(x.shape, factor))
    t = np.arange(factor + 1) / factor
    for c in itertools.product(*[range(n - 1) for n in x.shape]):
        block = np.zeros((factor + 1,) * N)
        for corner in itertools.product((0, 1), repeat=N):
            w = np.ones(())
            for k, b in enumerate(corner):
                wk = t if b else 1 - t
                w = np.multiply.outer(w, wk)
            block += w * x[tuple(ci + b for ci, b in zip(c, corner))]
        out[tuple(slice(ci * factor, ci * factor + factor + 1) for ci in c)] = block
    return out


def upsample_backprop(g, in_dims, factor):
    """UpsampleFilter::backprop (TopologyOptimizationFilter.hh:494-524): coarse node n gathers the fine nodes of
    its hat's support, weighted by prod_k (1 - |v_k - f n_k| / f)."""
    g = np.asarray(g, float)
    out = np.zeros(in_dims)
    for n in itertools.product(*[range(d) for d in in_dims]):
        lo = [nk * factor - (factor - 1) if nk > 0 else 0 for nk in n]
        hi = [nk * factor + (factor - 1) if nk < d - 1 else nk * factor for nk, d in zip(n, in_dims)]
        s = 0.0
        for v in itertools.product(*[range(a, b + 1) for a, b in zip(lo, hi)]):
            w = 1.0
            for vk, nk in zip(v, n):
                w *= 1 - abs(vk - nk * factor) / factor
            s += w * g[v]
        out[n] = s
    return out


def vertex_to_cell(x):
    """VertexToCellFilter::apply (TopologyOptimizationFilter.hh:545-560): each cell is 2^-N times the sum of its
    2^N corner vertices."""
    x = np.asarray(x, float)
    N = x.ndim
    out = np.zeros(tuple(n - 1 for n in x.shape))
    for corner in itertools.product((0, 1), repeat=N):
        out += x[tuple(slice(b, b + n - 1) for b, n in zip(corner, x.shape))]
    return out * 2.0 ** -N


def vertex_to_cell_backprop(g, in_dims):
    """VertexToCellFilter::backprop (TopologyOptimizationFilter.hh:563-583): each vertex is 2^-N times the sum of
    the cells it is a corner of."""
    g = np.asarray(g, float)
    N = g.ndim
    out = np.zeros(in_dims)
    for corner in itertools.product((0, 1), repeat=N):
        out[tuple(slice(b, b + n) for b, n in zip(corner, g.shape))] += g
    return out * 2.0 ** -N


def _smin(a, b):
    return 0.5 * (a + b - ((a - b) * (a - b) + LG_EPS) ** 0.5 + LG_EPS ** 0.5)


def _dsmin_dx1(a, b):
    return 0.5 * (1 - (a - b) * ((a - b) * (a - b) + LG_EPS) ** -0.5)


def _dsmin_dx2(a, b):
    return 0.5 * (1 + (a - b) * ((a - b) * (a - b) + LG_EPS) ** -0.5)


def _support(nx, x):
    """visitSupportingRegion at N = 2 (NDVector.hh:212-229): the voxel below, then x - 1 and x + 1 in bounds."""
    s = [x]
    if x > 0:
        s.append(x - 1)
    if x + 1 < nx:
        s.append(x + 1)
    return s


def _psum(v, nx, x, ylo):
    s = 0.0
    for xs in _support(nx, x):
        s += v[xs, ylo] ** LG_P
    return s


def langelaar(x):
    """LangelaarFilter::apply on a planar (nx, ny) grid (TopologyOptimizationFilter.hh:609-623): layer 0 copies
    the input; layer y is smin(x, smax(filtered supports in layer y - 1)), smax = (sum v^P)^(1/Q).
    Returns (filtered, smax) with smax zero on layer 0."""
    x = np.asarray(x, float)
    nx, ny = x.shape
    out = np.zeros_like(x)
    smax = np.zeros_like(x)
    out[:, 0] = x[:, 0]
    for y in range(1, ny):
        for i in range(nx):
            smax[i, y] = _psum(out, nx, i, y - 1) ** (1 / LG_Q)
            out[i, y] = _smin(x[i, y], smax[i, y])
    return out, smax


def langelaar_backprop(g, x, filtered, smax):
    """LangelaarFilter::backprop (TopologyOptimizationFilter.hh:625-662): multipliers from the top layer down,
    lambda_k = g_k + sum over the voxels i of the layer above that k supports (in visiting order) of
    lambda_i (dsmin/dx2(x_i, smax_i) dsmax_i/dfiltered_k); the result is lambda (layer 0) or
    lambda dsmin/dx1(x, smax)."""
    g = np.asarray(g, float)
    nx, ny = g.shape
    lam = np.zeros_like(g)
    for y in range(ny - 1, -1, -1):
        lam[:, y] = g[:, y]
        if y < ny - 1:
            for i in range(nx):
                total = _psum(filtered, nx, i, y)
                for k in _support(nx, i):
                    dsmax = LG_P * filtered[k, y] ** (LG_P - 1) / LG_Q * total ** (1 / LG_Q - 1)
                    lam[k, y] += lam[i, y + 1] * (_dsmin_dx2(x[i, y + 1], smax[i, y + 1]) * dsmax)
    out = lam.copy()
    out[:, 1:] = lam[:, 1:] * _dsmin_dx1(x[:, 1:], smax[:, 1:])
    return out


def am_filter_exact(x):
    """The additive-manufacturing filter that LangelaarFilter smooths: layer y = min(x, max(filtered supports))."""
    x = np.asarray(x, float)
    nx, ny = x.shape
    out = x.copy()
    for y in range(1, ny):
        for i in range(nx):
            out[i, y] = min(x[i, y], max(out[k, y - 1] for k in _support(nx, i)))
    return out
