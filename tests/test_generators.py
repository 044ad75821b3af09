"""CPU tests of the workload generators of the convection-diffusion (C2) and
random-conductivity (C3) configurations, against independent restatements
and closed forms, and of the binding's argument checks."""
from fractions import Fraction

import numpy as np
import pytest

import parcel_hip as ph
from problems import convdiff3d


def test_convdiff3d_equals_scaled_restatement():
    # problems.convdiff3d builds the same operator with diffusion 1/h^2 and
    # convection v/h; the library's is that times h^2
    for g, pe in ((6, 0.5), (5, 2.0)):
        rp, col, val = ph.convdiff3d(g, g, g, pe, (1.0, 0.5, 0.25))
        r2, c2, v2, _ = convdiff3d(g, peclet=pe)
        h = 1.0 / (g + 1)
        assert np.array_equal(rp, r2) and np.array_equal(col, c2)
        assert np.allclose(val, v2 * h * h, rtol=1e-13, atol=0)


def test_convdiff3d_c2_coefficients():
    g = 5
    rp, col, val = ph.convdiff3d(g, g, g)
    n = g ** 3
    A = np.zeros((n, n))
    for r in range(n):
        for e in range(rp[r] - 1, rp[r + 1] - 1):
            A[r, col[e] - 1] = val[e]
    # Pe 0.5 with direction (1, 0.5, 0.25): upwind -2, -1.5, -1.25, downwind -1
    r = 2 + g * (2 + g * 2)  # the interior cell (2, 2, 2)
    assert A[r, r] == 7.75
    assert (A[r, r - 1], A[r, r + 1]) == (-2.0, -1.0)
    assert (A[r, r - g], A[r, r + g]) == (-1.5, -1.0)
    assert (A[r, r - g * g], A[r, r + g * g]) == (-1.25, -1.0)
    assert np.all(np.diag(A) == 7.75)
    # columns of a row sorted ascending (the reference's b,s,w,c,e,n,t order)
    for q in range(n):
        c = col[rp[q] - 1: rp[q + 1] - 1]
        assert np.all(np.diff(c) > 0)
    # diagonally dominant M-matrix; nonsymmetric
    assert np.all(np.abs(A).sum(axis=1) - 2 * np.abs(np.diag(A)) <= 1e-15)
    assert not np.array_equal(A, A.T)


def test_convdiff3d_negative_velocity_mirrors():
    g = 4
    rp, col, v1 = ph.convdiff3d(g, g, g, 0.5, (1.0, 0.5, 0.25))
    _, _, v2 = ph.convdiff3d(g, g, g, 0.5, (-1.0, -0.5, -0.25))
    n = g ** 3
    A1, A2 = np.zeros((n, n)), np.zeros((n, n))
    for r in range(n):
        for e in range(rp[r] - 1, rp[r + 1] - 1):
            A1[r, col[e] - 1] = v1[e]
            A2[r, col[e] - 1] = v2[e]
    # reversing the velocity transposes the convective part
    assert np.array_equal(A1, A2.T)


def _diffusion3d_restated(g, kappa):
    """problem.f90:620-770 (make_poisson3D_adv_CRS) in plain Python, ghost =
    boundary cell, dense."""
    n = g ** 3
    k = kappa.reshape(g, g, g)  # [z, y, x]
    dx = dy = dz = 1.0 / g
    A = np.zeros((n, n))
    idx = lambda ix, iy, iz: ix + g * (iy + g * iz)
    kk = lambda ix, iy, iz, dflt: k[iz, iy, ix] if (0 <= ix < g and 0 <= iy < g and 0 <= iz < g) else dflt
    for iz in range(g):
        for iy in range(g):
            for ix in range(g):
                kc = k[iz, iy, ix]
                f = lambda kn, a, b, c: -(a * b / (c * 0.5 * (1.0 / kc + 1.0 / kn)))
                cw = f(kk(ix - 1, iy, iz, kc), dy, dz, dx)
                ce = f(kk(ix + 1, iy, iz, kc), dy, dz, dx)
                cs = f(kk(ix, iy - 1, iz, kc), dx, dz, dy)
                cn = f(kk(ix, iy + 1, iz, kc), dx, dz, dy)
                cb = f(kk(ix, iy, iz - 1, kc), dx, dy, dz)
                ct = f(kk(ix, iy, iz + 1, kc), dx, dy, dz)
                r = idx(ix, iy, iz)
                A[r, r] = -(cw + ce + cs + cn + cb + ct)
                for (jx, jy, jz, c) in ((ix - 1, iy, iz, cw), (ix + 1, iy, iz, ce), (ix, iy - 1, iz, cs),
                                        (ix, iy + 1, iz, cn), (ix, iy, iz - 1, cb), (ix, iy, iz + 1, ct)):
                    if 0 <= jx < g and 0 <= jy < g and 0 <= jz < g:
                        A[r, idx(jx, jy, jz)] = c
    return A


def test_diffusion3d_matches_restatement_bit_exact():
    g = 5
    kappa = ph.loguniform(g ** 3, 2, 1e-4, 1e4)
    rp, col, val = ph.diffusion3d(g, g, g, kappa)
    n = g ** 3
    A = np.zeros((n, n))
    for r in range(n):
        for e in range(rp[r] - 1, rp[r + 1] - 1):
            A[r, col[e] - 1] = val[e]
    assert np.array_equal(A, _diffusion3d_restated(g, kappa))
    assert np.array_equal(A, A.T)
    assert np.all(np.linalg.eigvalsh(A) > 0)


def test_diffusion3d_constant_kappa_is_scaled_poisson():
    g = 8
    rp, col, val = ph.diffusion3d(g, g, g, np.ones(g ** 3))
    r2, c2, v2 = ph.poisson3d_spd(g, g, g)
    assert np.array_equal(rp, r2) and np.array_equal(col, c2)
    # face = dy*dz/dx = 1/g; boundary faces (ghost = own kappa) stay on the diagonal
    assert np.array_equal(val, v2 / g)


def test_loguniform_range_and_moments():
    k = ph.loguniform(200000, 2, 1e-4, 1e4)
    assert k.min() >= 1e-4 * (1 - 1e-15) and k.max() < 1e4
    lk = np.log10(k)
    assert abs(lk.mean()) < 0.05 and abs(lk.std() - 8 / np.sqrt(12)) < 0.02
    u = ph.uniform(200000, 2, 0.0, 1.0)
    assert np.allclose(k, np.exp(np.log(1e-4) + (np.log(1e4) - np.log(1e-4)) * u), rtol=1e-15, atol=0)


def test_spmv_dd_is_exactly_rounded_sum():
    g = 4
    kappa = ph.loguniform(g ** 3, 5, 1e-4, 1e4)
    rp, col, val = ph.diffusion3d(g, g, g, kappa)
    x = ph.uniform(g ** 3, 6)
    bh, bl = ph.spmv_dd(rp, col, val, x)
    for r in range(g ** 3):
        exact = sum(Fraction(val[e]) * Fraction(x[col[e] - 1]) for e in range(rp[r] - 1, rp[r + 1] - 1))
        got = Fraction(bh[r]) + Fraction(bl[r])
        scale = max(abs(Fraction(val[e]) * Fraction(x[col[e] - 1])) for e in range(rp[r] - 1, rp[r + 1] - 1))
        assert abs(got - exact) <= scale * Fraction(1, 2 ** 100)
        assert bh[r] == float(got)  # hi is the rounded value of the pair


def test_binding_checks_dtypes_and_sizes():
    rp, col, val = ph.poisson2d(4, 4, -1.0)
    n = 16
    b = np.ones(n)
    with pytest.raises(ValueError):
        ph.dcg(np.zeros(n, np.float32), b, rp, col, val)  # x must be float64 (written in place)
    with pytest.raises(ValueError):
        ph.dcg(np.zeros(n + 1), b, rp, col, val)  # x length
    with pytest.raises(ValueError):
        ph.dcg(np.zeros(n), np.ones(n - 1), rp, col, val)  # b length
    with pytest.raises(ValueError):
        ph.dcg(np.zeros(n), b, rp, col, val[:-1])  # val length
    with pytest.raises(ValueError):
        ph.dgmres(np.zeros(2 * n)[::2], b, rp, col, val)  # x not contiguous
    # inputs of other integer / float types are converted, not misread: the
    # converted call only reaches the library, which then reports no device
    with pytest.raises(ph.ParcelError):
        ph.dcg(np.zeros(n), b.astype(np.float32), rp.astype(np.int64), col.astype(np.int64), val)


def test_crs_to_dia_round_trip():
    import scipy.sparse as sp
    for rp, col, val in (ph.poisson3d_parcel(5, 4, 3), convdiff3d(4, peclet=1.0)[:3], ph.poisson2d(7, 3, -1.0)):
        n = rp.shape[0] - 1
        off, dia = ph.crs_to_dia(rp, col, val)
        assert np.all(np.diff(off) > 0)
        D = np.zeros((n, n))
        for j, o in enumerate(off):
            for i in range(n):
                if 0 <= i + o < n:
                    D[i, i + o] = dia[j * n + i]
                else:
                    assert dia[j * n + i] == 0.0
        A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n)).toarray()
        assert np.array_equal(D, A)
