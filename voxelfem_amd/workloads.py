"""Seeded synthetic workloads of the solve-and-sensitivity benchmarks.

Each workload is a voxel grid with unit voxels (domain [0,nx]x[0,ny]x[0,nz], the reference's default
when no bounding box is given, TensorProductSimulator.hh:282-285), isotropic material E=1, nu=0.3,
SIMP with E0=1, Emin=1e-9, p=3, and boundary conditions expressed in the reference's own `.bc`
region language (box% fractions of the domain) so that they reach the simulator through
``applyDisplacementsAndLoadsFromFile`` exactly like the reference drivers' inputs.

Densities are drawn in C (element) order with ``numpy.random.default_rng(seed)``.
"""
import json
import os
import tempfile

import numpy as np

CANTILEVER_BC = {
    # x = 0 clamped; unit downward (-z) load spread over the line x = L, z = H/2 (all y)
    "regions": [
        {"type": "dirichlet", "value": [0, 0, 0],
         "box%": {"minCorner": [-1e-4, -1e-4, -1e-4], "maxCorner": [1e-4, 1.0001, 1.0001]}},
        {"type": "force", "value": [0, 0, -1],
         "box%": {"minCorner": [0.9999, -1e-4, 0.4999], "maxCorner": [1.0001, 1.0001, 0.5001]}},
    ]
}

MBB_BC = {
    # roller (u_y, u_z) on the bottom edge at x = L; unit -y load at the top of the symmetry plane x = 0
    # (u_x = 0 there is added with applySymmetryConditions)
    "regions": [
        {"type": "dirichletyz", "value": [0, 0, 0],
         "box%": {"minCorner": [0.9999, -1e-4, -1e-4], "maxCorner": [1.0001, 1e-4, 1.0001]}},
        {"type": "force", "value": [0, -1, 0],
         "box%": {"minCorner": [-1e-4, 0.9999, -1e-4], "maxCorner": [1e-4, 1.0001, 1.0001]}},
    ]
}

MATERIAL = {"type": "isotropic_material", "dim": 3, "density": 1.0, "young": 1.0, "poisson": 0.3}

# planar (2D) workloads: the region language is the same; a planar grid reads the first two entries of
# every corner and value
CANTILEVER2D_BC = {
    # x = 0 clamped; unit downward load at the lower right corner (x = L, y = 0)
    "regions": [
        {"type": "dirichlet", "value": [0, 0, 0],
         "box%": {"minCorner": [-1e-4, -1e-4, 0], "maxCorner": [1e-4, 1.0001, 0]}},
        {"type": "force", "value": [0, -1, 0],
         "box%": {"minCorner": [0.9999, -1e-4, 0], "maxCorner": [1.0001, 1e-4, 0]}},
    ]
}

MBB2D_BC = {
    # pinned lower left corner, roller (u_y) at the lower right corner, unit downward load at the top middle
    "regions": [
        {"type": "dirichlet", "value": [0, 0, 0],
         "box%": {"minCorner": [-1e-4, -1e-4, 0], "maxCorner": [1e-4, 1e-4, 0]}},
        {"type": "dirichlety", "value": [0, 0, 0],
         "box%": {"minCorner": [0.9999, -1e-4, 0], "maxCorner": [1.0001, 1e-4, 0]}},
        {"type": "force", "value": [0, -1, 0],
         "box%": {"minCorner": [0.4999, 0.9999, 0], "maxCorner": [0.5001, 1.0001, 0]}},
    ]
}

GRIDS = {
    "cantilever_64x32x32": (64, 32, 32),
    "cantilever_256x128x128": (256, 128, 128),
    "mbb_128x64x64": (128, 64, 64),
    "lbl_160x80x80": (160, 80, 80),
    "cantilever_384x192x192": (384, 192, 192),
}


def noise_density(ne, seed, base=0.3, amp=0.05):
    """rho = clamp(base + U(-amp, amp), 1e-3, 1)"""
    rng = np.random.default_rng(seed)
    n = int(np.prod(ne))
    return np.clip(base + rng.uniform(-amp, amp, n), 1e-3, 1.0)


def smoothed_density(ne, seed, sigma=3.0, mean=0.3, std=0.1):
    """U(0,1) voxel noise, Gaussian-smoothed (sigma voxels, reflect, truncate 4), rescaled affinely to
    the given mean/std, clamped to [1e-3, 1], shifted once to restore the mean and clamped again."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, int(np.prod(ne))).reshape(ne)
    g = gaussian_filter(x, sigma=sigma, mode="reflect", truncate=4.0)
    z = (g - g.mean()) / g.std() * std + mean
    z = np.clip(z, 1e-3, 1.0)
    z = np.clip(z + (mean - z.mean()), 1e-3, 1.0)
    return z.ravel()


def write_json(obj, directory=None, name="bc.json"):
    directory = directory or tempfile.mkdtemp(prefix="vf_wl_")
    path = os.path.join(directory, name)
    with open(path, "w") as f:
        json.dump(obj, f)
    return path


def configure(pv, ne, bc, rho, directory=None, Emin=1e-9, symmetry_x=False):
    """Create a pyVoxelFEM simulator with the standard material/law, `bc` (a .bc dict), optional
    u_x = 0 symmetry on the x = 0 face, and densities."""
    tps = pv.TensorProductSimulator([1, 1, 1], [[0.0, 0.0, 0.0], [float(n) for n in ne]], list(ne))
    d = directory or tempfile.mkdtemp(prefix="vf_wl_")
    tps.readMaterial(write_json(MATERIAL, d, "material.json"))
    tps.E_0 = 1.0
    tps.E_min = Emin
    tps.gamma = 3.0
    if bc is not None:
        tps.applyDisplacementsAndLoadsFromFile(write_json(bc, d))
    if symmetry_x:
        tps.applySymmetryConditions([True, False, False], [False, False, False])
    if rho is not None:
        tps.setDensities(np.asarray(rho, dtype=np.float64))
    return tps


def configure2d(pv, ne, bc, rho, directory=None, Emin=1e-9, law=None):
    """A planar ([1, 1]) pyVoxelFEM simulator on [0, nx] x [0, ny] with the standard material and law."""
    tps = pv.TensorProductSimulator([1, 1], [[0.0, 0.0], [float(n) for n in ne]], list(ne))
    d = directory or tempfile.mkdtemp(prefix="vf_wl_")
    tps.readMaterial(write_json(MATERIAL, d, "material.json"))
    tps.E_0 = 1.0
    tps.E_min = Emin
    tps.gamma = 3.0
    if law is not None:
        tps.interpolationLaw = law
    if bc is not None:
        tps.applyDisplacementsAndLoadsFromFile(write_json(bc, d))
    if rho is not None:
        tps.setDensities(np.asarray(rho, dtype=np.float64))
    return tps


def cantilever2d_mask(ne):
    X = _node_positions(ne)
    return np.where(X[:, 0] == 0, 3, 0).astype(np.uint8)


def cantilever2d_load(ne):
    X = _node_positions(ne)
    f = np.zeros((len(X), 2))
    f[(X[:, 0] == ne[0]) & (X[:, 1] == 0), 1] = -1.0
    return f


def mbb2d_mask(ne):
    X = _node_positions(ne)
    m = np.zeros(len(X), np.uint8)
    m[(X[:, 0] == 0) & (X[:, 1] == 0)] = 3
    m[(X[:, 0] == ne[0]) & (X[:, 1] == 0)] = 2
    return m


def mbb2d_load(ne):
    X = _node_positions(ne)
    f = np.zeros((len(X), 2))
    top = (X[:, 0] == ne[0] / 2) & (X[:, 1] == ne[1])
    f[top, 1] = -1.0 / top.sum()
    return f


def _node_positions(ne):
    nn = [n + 1 for n in ne]
    g = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in nn], indexing="ij")
    return np.stack([a.ravel() for a in g], axis=1)


def cantilever_load(ne):
    """(nn,3) load of CANTILEVER_BC, built directly (for the oracle)."""
    X = _node_positions(ne)
    line = (X[:, 0] == ne[0]) & (X[:, 2] == ne[2] / 2)
    f = np.zeros((len(X), 3))
    f[line, 2] = -1.0 / line.sum()
    return f


def cantilever_mask(ne):
    X = _node_positions(ne)
    return np.where(X[:, 0] == 0, 7, 0).astype(np.uint8)


def mbb_mask(ne):
    """u_x = 0 on x = 0 (symmetry) and (u_y, u_z) = 0 on the edge x = L, y = 0 (roller)."""
    X = _node_positions(ne)
    m = np.where(X[:, 0] == 0, 1, 0).astype(np.uint8)
    m[(X[:, 0] == ne[0]) & (X[:, 1] == 0)] |= 6
    return m


def mbb_load(ne):
    X = _node_positions(ne)
    top = (X[:, 0] == 0) & (X[:, 1] == ne[1])
    f = np.zeros((len(X), 3))
    f[top, 1] = -1.0 / top.sum()
    return f
