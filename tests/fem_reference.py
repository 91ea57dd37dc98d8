"""Independent brute-force constructions used to pin the oracle on tiny grids.

Everything here is assembled from the mathematical definitions (sparse global matrices via
scipy, trilinear interpolation on nested grids), not from the oracle's code paths.
"""
import numpy as np
import scipy.sparse as sp


def node_index(nn, ix, iy, iz):
    return (ix * nn[1] + iy) * nn[2] + iz


def elem_nodes(ne):
    """(nelems, 8) global node indices, local order l = 4i+2j+k."""
    nn = [n + 1 for n in ne]
    ex, ey, ez = np.meshgrid(np.arange(ne[0]), np.arange(ne[1]), np.arange(ne[2]), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    cols = []
    for l in range(8):
        i, j, k = (l >> 2) & 1, (l >> 1) & 1, l & 1
        cols.append(node_index(nn, ex + i, ey + j, ez + k))
    return np.stack(cols, axis=1)


def assemble_K(ne, K0, E):
    """Global stiffness with DoF order 3n+c from element matrices E_e*K0."""
    en = elem_nodes(ne)
    dofs = (3 * en[:, :, None] + np.arange(3)[None, None, :]).reshape(len(en), 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = (E[:, None, None] * K0[None]).reshape(-1)
    nn = int(np.prod([n + 1 for n in ne]))
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * nn, 3 * nn))


def assemble_from_elements(ne, Kes):
    en = elem_nodes(ne)
    dofs = (3 * en[:, :, None] + np.arange(3)[None, None, :]).reshape(len(en), 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    nn = int(np.prod([n + 1 for n in ne]))
    return sp.csr_matrix((Kes.reshape(-1), (rows, cols)), shape=(3 * nn, 3 * nn))


def prolongation_1d(nc):
    """(2*(nc-1)+1, nc) linear interpolation between nested 1D node grids."""
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
    """Scalar trilinear prolongation (node ordering C, z fastest), then expanded to 3 comps."""
    Px, Py, Pz = (sp.csr_matrix(prolongation_1d(n)) for n in nn_coarse)
    P = sp.kron(sp.kron(Px, Py), Pz).tocsr()
    return sp.kron(P, sp.identity(3)).tocsr()


def to_dof(u):
    """(nn,3) -> DoF vector 3n+c"""
    return np.asarray(u).reshape(-1)


def from_dof(v):
    return np.asarray(v).reshape(-1, 3)


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def element_node_positions(h):
    X = np.zeros((8, 3))
    for l in range(8):
        X[l] = [((l >> 2) & 1) * h[0], ((l >> 1) & 1) * h[1], (l & 1) * h[2]]
    return X


def block_gauss_seidel(K, u, b, colors, dmask, forward=True):
    """Sequential multicolour block Gauss-Seidel straight from the definition, on a global
    sparse matrix K (DoF order 3n+c). colors: list of node-index arrays in sweep order."""
    K = K.tocsr()
    u = u.copy()
    for nodes in colors:
        for n in nodes:
            m = dmask[n]
            if m == 7:
                continue
            rows = slice(3 * n, 3 * n + 3)
            Kr = K[rows, :]
            res = b[rows] - Kr @ u
            M = Kr[:, 3 * n:3 * n + 3].toarray()
            if m:
                ud = np.zeros(3)
                order = range(3) if forward else range(2, -1, -1)
                for i in order:
                    free = 0.0 if (m >> i) & 1 else 1.0
                    ud[i] = (res[i] - M[i] @ ud) * free / M[i, i]
                u[rows] += ud
            else:
                u[rows] += np.linalg.solve(M, res)
    return u
