"""Brute-force planar (2D) constructions used to pin the planar oracle on tiny grids.

Assembled from the mathematical definitions (scipy sparse global matrices, bilinear interpolation between
nested grids), not from the oracle's code paths. DoF order 2 n + c, node n = ix * nny + iy, local element
node l = 2 i + j for the corner (i, j).
"""
import numpy as np
import scipy.sparse as sp


def elem_nodes(ne):
    nny = ne[1] + 1
    ex, ey = np.meshgrid(np.arange(ne[0]), np.arange(ne[1]), indexing="ij")
    ex, ey = ex.ravel(), ey.ravel()
    return np.stack([(ex + ((l >> 1) & 1)) * nny + ey + (l & 1) for l in range(4)], axis=1)


def assemble_from_elements(ne, Kes):
    en = elem_nodes(ne)
    dofs = (2 * en[:, :, None] + np.arange(2)[None, None, :]).reshape(len(en), 8)
    rows = np.repeat(dofs, 8, axis=1).ravel()
    cols = np.tile(dofs, (1, 8)).ravel()
    nn = (ne[0] + 1) * (ne[1] + 1)
    return sp.csr_matrix((np.asarray(Kes).reshape(-1), (rows, cols)), shape=(2 * nn, 2 * nn))


def assemble_K(ne, K0, E):
    return assemble_from_elements(ne, np.asarray(E)[:, None, None] * np.asarray(K0)[None])


def prolongation_1d(nc):
    nf = 2 * (nc - 1) + 1
    P = np.zeros((nf, nc))
    for i in range(nf):
        if i % 2 == 0:
            P[i, i // 2] = 1.0
        else:
            P[i, (i - 1) // 2] = 0.5
            P[i, (i + 1) // 2] = 0.5
    return P


def prolongation(nn_coarse):
    Px, Py = (sp.csr_matrix(prolongation_1d(n)) for n in nn_coarse)
    return sp.kron(sp.kron(Px, Py), sp.identity(2)).tocsr()


def to_dof(u):
    return np.asarray(u).reshape(-1)


def from_dof(v):
    return np.asarray(v).reshape(-1, 2)


def plane_stress_lame(E, nu):
    """Lame constants of the plane-stress tensor: sigma = lam* tr(eps) I + 2 mu eps."""
    return E * nu / (1 - nu * nu), E / (2 * (1 + nu))


def top88_KE(E, nu):
    """Closed-form plane-stress bilinear quad stiffness of the unit square (Andreassen et al., "Efficient
    topology optimization in MATLAB using 88 lines of code", 2011), nodes counter-clockwise from the
    lower-left corner: (0,0), (1,0), (1,1), (0,1), DoFs (u, v) per node."""
    k = np.array([1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12, -1 / 8 + 3 * nu / 8,
                  -1 / 4 + nu / 12, -1 / 8 - nu / 8, nu / 6, 1 / 8 - 3 * nu / 8])
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2], [2, 7, 0, 5, 6, 3, 4, 1], [3, 6, 5, 0, 7, 2, 1, 4],
           [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6], [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    return E / (1 - nu * nu) * k[np.array(idx)]


def block_gauss_seidel(K, u, b, colors, dmask, forward=True):
    """Sequential multicolour block Gauss-Seidel from the definition on a global sparse K (DoF 2 n + c)."""
    K = K.tocsr()
    u = u.copy()
    for nodes in colors:
        for n in nodes:
            m = dmask[n]
            if m == 3:
                continue
            rows = slice(2 * n, 2 * n + 2)
            Kr = K[rows, :]
            res = b[rows] - Kr @ u
            M = Kr[:, 2 * n:2 * n + 2].toarray()
            if m:
                ud = np.zeros(2)
                for i in (range(2) if forward else range(1, -1, -1)):
                    free = 0.0 if (m >> i) & 1 else 1.0
                    ud[i] = (res[i] - M[i] @ ud) * free / M[i, i]
                u[rows] += ud
            else:
                u[rows] += np.linalg.solve(M, res)
    return u
