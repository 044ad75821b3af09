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


def ddm_box_split(n):
    """One direction of the GPU build's box partition (init_DILU_DDM with n/4
    boxes, src/src_DDM/precon_ddm_GPU.f90:399-470): box starts and the end."""
    div = max(n // 4, 1)
    loc = n // div + (1 if n % div else 0)
    edges = [loc * i for i in range(div) if min(n, loc * (i + 1)) > loc * i]
    return edges + [n]


def ddm_box_crs(d):
    """The stencil problem d (parcel_hip.poisson3d_ddm) as 1-based CSR in
    box-major numbering: (rowptr, col, val, perm, starts), perm[r] = padded
    index of row r, starts = 0-based first row of every box (and n)."""
    nx, ny, nz = d["grid"]
    mg = d["margin"]
    mx, my = nx + 2 * mg, ny + 2 * mg
    m = d["m"]
    val = d["val_dia"].reshape(7, m)
    ioff = d["ioff_dia"]
    ex, ey, ez = ddm_box_split(nx), ddm_box_split(ny), ddm_box_split(nz)
    perm, starts = [], []
    for bz in range(len(ez) - 1):
        for by in range(len(ey) - 1):
            for bx in range(len(ex) - 1):
                starts.append(len(perm))
                for jz in range(ez[bz], ez[bz + 1]):
                    for jy in range(ey[by], ey[by + 1]):
                        for jx in range(ex[bx], ex[bx + 1]):
                            perm.append((jx + mg) + mx * (jy + mg) + mx * my * (jz + mg))
    starts.append(len(perm))
    perm = np.array(perm)
    pos = -np.ones(m, np.int64)
    pos[perm] = np.arange(perm.size)
    rows, cols, vals = [], [], []
    for r, j in enumerate(perm):
        ent = []
        for ip in range(7):
            v = val[ip, j]
            jn = j + ioff[ip]
            if pos[jn] < 0 or (v == 0.0 and jn != j):
                continue
            ent.append((pos[jn], v))
        for c, v in sorted(ent):
            rows.append(r)
            cols.append(c)
            vals.append(v)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(perm.size, perm.size))
    A.sort_indices()
    return (A.indptr.astype(np.int32) + 1, A.indices.astype(np.int32) + 1, A.data.copy(), perm,
            np.array(starts, np.int32))


def ddm_ras(d, addL):
    """Restricted additive Schwarz of the GPU build (init_AS_DDM,
    src/src_DDM/precon_ddm_GPU.f90:370-643, one rank): the boxes of
    ddm_box_split extended by addL cells on every side facing another box;
    expanded row q = a cell of an extended box (boxes in box order, cells in
    z, y, x order).  Returns (src, dst, rp_e, col_e, val_e, starts_e) in the
    box-major numbering of ddm_box_crs: src[q] = the row gathered, dst[q] =
    the row written back (-1 for overlap cells), the expanded matrix keeps
    the couplings inside each extended box (1-based CSR)."""
    nx, ny, nz = d["grid"]
    mg = d["margin"]
    mx, my = nx + 2 * mg, ny + 2 * mg
    m = d["m"]
    val = d["val_dia"].reshape(7, m)
    ioff = d["ioff_dia"]
    ex, ey, ez = ddm_box_split(nx), ddm_box_split(ny), ddm_box_split(nz)
    _, _, _, perm, _ = ddm_box_crs(d)
    pos = -np.ones(m, np.int64)
    pos[perm] = np.arange(perm.size)
    lo = lambda e, k: e[k] - addL if k > 0 else e[k]
    hi = lambda e, k: e[k + 1] + addL if k < len(e) - 2 else e[k + 1]
    src, dst, starts, rows, cols, vals = [], [], [], [], [], []
    for bz in range(len(ez) - 1):
        for by in range(len(ey) - 1):
            for bx in range(len(ex) - 1):
                base = len(src)
                starts.append(base)
                cells = []
                for jz in range(lo(ez, bz), hi(ez, bz)):
                    for jy in range(lo(ey, by), hi(ey, by)):
                        for jx in range(lo(ex, bx), hi(ex, bx)):
                            j = (jx + mg) + mx * (jy + mg) + mx * my * (jz + mg)
                            own = (ex[bx] <= jx < ex[bx + 1] and ey[by] <= jy < ey[by + 1] and
                                   ez[bz] <= jz < ez[bz + 1])
                            src.append(pos[j])
                            dst.append(pos[j] if own else -1)
                            cells.append(j)
                local = {j: base + q for q, j in enumerate(cells)}
                for q, j in enumerate(cells):
                    ent = []
                    for ip in range(7):
                        jn = j + ioff[ip]
                        v = val[ip, j]
                        if jn not in local or (v == 0.0 and jn != j):
                            continue
                        ent.append((local[jn], v))
                    for c, v in sorted(ent):
                        rows.append(base + q)
                        cols.append(c)
                        vals.append(v)
    ne = len(src)
    starts.append(ne)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(ne, ne))
    A.sort_indices()
    return (np.array(src, np.int32), np.array(dst, np.int32), A.indptr.astype(np.int32) + 1,
            A.indices.astype(np.int32) + 1, A.data.copy(), np.array(starts, np.int32))

