"""Small test problems shared by the CPU and GPU tests (1-based CSR)."""
import numpy as np
import scipy.sparse as sp


def csr_to_scipy(rp, col, val):
    n = rp.shape[0] - 1
    return sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))


def convdiff3d(g, peclet=0.5, seed=None):
    """Small nonsymmetric 7-point upwind convection-diffusion operator (1-based CSR)."""
    n = g ** 3
    h = 1.0 / (g + 1)
    vel = np.array([1.0, 0.5, 0.25])
    vel = vel * (2.0 * peclet / (h * np.abs(vel).max()))
    rows, cols, vals = [], [], []
    idx = lambda i, j, k: i + g * (j + g * k)
    for k in range(g):
        for j in range(g):
            for i in range(g):
                r = idx(i, j, k)
                diag = 6.0 / h ** 2 + np.abs(vel).sum() / h
                for d, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    for sgn in (-1, 1):
                        ii, jj, kk = i + sgn * di, j + sgn * dj, k + sgn * dk
                        if 0 <= ii < g and 0 <= jj < g and 0 <= kk < g:
                            c = -1.0 / h ** 2
                            if sgn == -1:
                                c -= vel[d] / h  # upwind for positive velocity
                            rows.append(r); cols.append(idx(ii, jj, kk)); vals.append(c)
                rows.append(r); cols.append(r); vals.append(diag)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return A.indptr.astype(np.int32) + 1, A.indices.astype(np.int32) + 1, A.data.astype(np.float64), A
