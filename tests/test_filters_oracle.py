"""Pins of the numpy restatement of UpsampleFilter, VertexToCellFilter and LangelaarFilter (oracle/filters.py)
against independent routines and mathematical facts; CPU only."""
import itertools

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator
from scipy.signal import convolve

from oracle import filters as F


def _dense(fn, in_shape):
    """Matrix of a linear map on grids of in_shape, column by column."""
    n = int(np.prod(in_shape))
    cols = []
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1
        cols.append(np.asarray(fn(e.reshape(in_shape))).ravel())
    return np.array(cols).T


@pytest.mark.parametrize("shape,factor", [((3, 4), 2), ((2, 5), 3), ((4, 3), 1), ((3, 2, 3), 2), ((2, 3, 2), 3)])
def test_upsample_matches_linear_interpolation(shape, factor):
    rng = np.random.default_rng(sum(shape) + factor)
    x = rng.random(shape)
    out = F.upsample(x, factor)
    assert out.shape == F.upsample_dims(shape, factor)
    grids = [np.arange(n, dtype=float) for n in shape]
    pts = np.stack(np.meshgrid(*[np.arange(m) / factor for m in out.shape], indexing="ij"), axis=-1)
    ref = RegularGridInterpolator(grids, x, method="linear")(pts.reshape(-1, len(shape))).reshape(out.shape)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize("shape,factor", [((3, 4), 3), ((3, 3, 2), 2)])
def test_upsample_reproduces_multilinear_fields(shape, factor):
    """A multilinear field sampled on the coarse nodes is reproduced exactly on the fine nodes."""
    coef = np.array([0.3, -1.2, 0.7, 2.1, -0.4, 0.9, 1.3, -0.6])

    def field(*c):
        terms = [np.ones_like(c[0])]
        for r in range(1, len(c) + 1):
            for sub in itertools.combinations(range(len(c)), r):
                t = np.ones_like(c[0])
                for k in sub:
                    t = t * c[k]
                terms.append(t)
        return sum(a * t for a, t in zip(coef, terms))

    coarse = field(*np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij"))
    fine_dims = F.upsample_dims(shape, factor)
    fine = field(*np.meshgrid(*[np.arange(m) / factor for m in fine_dims], indexing="ij"))
    np.testing.assert_allclose(F.upsample(coarse, factor), fine, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("shape,factor", [((3, 4), 2), ((2, 3), 3), ((2, 3, 2), 2)])
def test_upsample_backprop_is_transpose(shape, factor):
    A = _dense(lambda e: F.upsample(e, factor), shape)
    out_shape = F.upsample_dims(shape, factor)
    B = _dense(lambda e: F.upsample_backprop(e, shape, factor), out_shape)
    np.testing.assert_allclose(B, A.T, rtol=0, atol=1e-15)


@pytest.mark.parametrize("shape", [(3, 4), (5, 2), (3, 4, 2), (2, 2, 2)])
def test_vertex_to_cell_is_box_average(shape):
    rng = np.random.default_rng(len(shape) * 7 + shape[0])
    x = rng.random(shape)
    ref = convolve(x, np.full((2,) * len(shape), 2.0 ** -len(shape)), mode="valid")
    np.testing.assert_allclose(F.vertex_to_cell(x), ref, rtol=1e-14, atol=0)


def test_vertex_to_cell_samples_linear_fields_at_cell_centres():
    shape = (4, 5, 3)
    g = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    x = 0.2 + 1.5 * g[0] - 0.7 * g[1] + 0.3 * g[2]
    c = np.meshgrid(*[np.arange(n - 1) + 0.5 for n in shape], indexing="ij")
    np.testing.assert_allclose(F.vertex_to_cell(x), 0.2 + 1.5 * c[0] - 0.7 * c[1] + 0.3 * c[2], rtol=1e-14)


@pytest.mark.parametrize("shape", [(3, 4), (2, 3, 3)])
def test_vertex_to_cell_backprop_is_transpose(shape):
    A = _dense(F.vertex_to_cell, shape)
    B = _dense(lambda e: F.vertex_to_cell_backprop(e, shape), tuple(n - 1 for n in shape))
    np.testing.assert_array_equal(B, A.T)


def test_langelaar_tracks_exact_am_filter_on_an_overhang():
    """A pillar at x = 3 carrying a ledge: the ledge is kept over the pillar and its two neighbours, and removed
    beyond them, as the min/max filter prescribes; the smoothing moves values by a few per cent at most."""
    x = np.full((8, 7), 0.0)
    x[3, :] = 1.0
    x[1:7, 4] = 1.0
    x[0:8, 6] = 0.6
    out, smax = F.langelaar(x)
    ex = F.am_filter_exact(x)
    assert ex[2, 4] == 1 and ex[4, 4] == 1 and ex[1, 4] == 0 and ex[6, 4] == 0
    np.testing.assert_allclose(out, ex, atol=0.06)
    assert np.all(smax[:, 0] == 0)
    np.testing.assert_array_equal(out[:, 0], x[:, 0])


def test_langelaar_single_column_closed_form():
    """With one column only the voxel below supports: smax = v^(P/Q)."""
    x = np.array([[0.9, 0.8, 0.95, 0.3, 0.7]])
    out, smax = F.langelaar(x)
    v = x[0, 0]
    for y in range(1, x.shape[1]):
        m = v ** (F.LG_P / F.LG_Q)
        assert abs(smax[0, y] - m) < 1e-14 * max(1, m)
        a, b = x[0, y], m
        v = 0.5 * (a + b - np.sqrt((a - b) ** 2 + 1e-4) + 1e-2)
        assert abs(out[0, y] - v) < 1e-14


@pytest.mark.parametrize("shape", [(5, 4), (1, 6), (7, 3)])
def test_langelaar_backprop_matches_central_differences(shape):
    rng = np.random.default_rng(shape[0] * 10 + shape[1])
    x = 0.3 + 0.6 * rng.random(shape)
    w = rng.standard_normal(shape)

    def J(v):
        return float(np.sum(w * F.langelaar(v)[0]))

    filtered, smax = F.langelaar(x)
    g = F.langelaar_backprop(w, x, filtered, smax)
    h = 1e-6
    fd = np.zeros(shape)
    for idx in np.ndindex(shape):
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd[idx] = (J(xp) - J(xm)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-7)
