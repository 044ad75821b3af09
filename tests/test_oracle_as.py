"""CPU tests pinning the oracle's additive Schwarz preconditioner of the CRS
path (AdditiveSchwarz_crs at one rank, src/matrix.f90:3809-4375,4749-4858,
5804-6117): closed forms where the local D-ILU solves are exact (a
tridiagonal matrix has no fill, so its D-ILU is its LU), the degenerate
cases that the reference turns into block Jacobi, and a planted overlap
mistake that must fail."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc


def tridiag(n, seed=0):
    rng = np.random.default_rng(seed)
    lo = -rng.uniform(0.5, 1.0, n - 1)
    up = -rng.uniform(0.5, 1.0, n - 1)
    d = 2.5 + rng.uniform(0.0, 1.0, n)
    A = sp.diags([lo, d, up], [-1, 0, 1], format="csr")
    A.sort_indices()
    return (A.indptr.astype(np.int32) + 1, A.indices.astype(np.int32) + 1, A.data.copy(), A.toarray())


def blocks(n, nb):
    out = []
    for t in range(nb):
        lo, hi = orc.divide_region(t, nb, n)
        out.append((lo - 1, hi))  # 0-based half-open
    return out


def expected(Ad, r, nb, addL, flag):
    """sum over the extended blocks of their exact local solves, by mode."""
    n = Ad.shape[0]
    s = np.zeros(n)
    for lo, hi in blocks(n, nb):
        a, b = max(lo - addL, 0), min(hi + addL, n)
        rin = r[a:b].copy()
        if flag in (3, 4):  # INTERPOLATE / NONE: the overlap's residual is zero
            rin[: lo - a] = 0.0
            rin[hi - a:] = 0.0
        z = np.linalg.solve(Ad[a:b, a:b], rin)
        if flag in (2, 4):  # RESTRICT / NONE: owned rows only
            s[lo:hi] = z[lo - a: hi - a]
        else:  # BASIC / INTERPOLATE: overlaps add up
            s[a:b] += z
    return s


@pytest.mark.parametrize("flag", [1, 2, 3, 4])
@pytest.mark.parametrize("n,nb,addL", [(40, 4, 2), (37, 3, 5), (64, 8, 4), (50, 5, 1)])
def test_as_exact_local_solves(n, nb, addL, flag):
    rp, col, val, Ad = tridiag(n, seed=n + addL)
    r = np.random.default_rng(1).standard_normal(n)
    for ilu in (1, 2):
        M = orc.AsPrecon(rp, col, val, ilu_method=ilu, nblocks=nb, addL=addL, iflagAS=flag)
        got = M.apply(r)
        want = expected(Ad, r, nb, addL, flag)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-13), (ilu, np.abs(got - want).max())


def test_as_planted_overlap_mistake_fails():
    # extending by addL - 1 instead of addL must not match
    n, nb, addL = 40, 4, 3
    rp, col, val, Ad = tridiag(n, seed=3)
    r = np.random.default_rng(2).standard_normal(n)
    got = orc.AsPrecon(rp, col, val, nblocks=nb, addL=addL, iflagAS=1).apply(r)
    assert not np.allclose(got, expected(Ad, r, nb, addL - 1, 1), rtol=1e-6)
    assert not np.allclose(got, expected(Ad, r, nb, addL, 2), rtol=1e-6)


def poisson2d(g):
    import parcel_hip as ph
    return ph.poisson2d(g, g, -1.0)


@pytest.mark.parametrize("ilu", [1, 2])
def test_as_degenerates_to_block_jacobi(ilu):
    rp, col, val = poisson2d(20)
    n = rp.shape[0] - 1
    r = np.random.default_rng(4).standard_normal(n)
    bj = orc.Precon(rp, col, val, 2, ilu_method=ilu, nblocks=4).apply(r)
    # addL = 0, and an overlap wider than half a block (100 rows): block Jacobi D-ILU
    for addL in (0, 51):
        for flag in (1, 2, 3, 4):
            got = orc.AsPrecon(rp, col, val, ilu_method=ilu, nblocks=4, addL=addL, iflagAS=flag).apply(r)
            assert np.array_equal(got, bj), (addL, flag)
    # an overlap of exactly half a block is still additive Schwarz
    got = orc.AsPrecon(rp, col, val, ilu_method=ilu, nblocks=4, addL=50, iflagAS=2).apply(r)
    assert not np.array_equal(got, bj)


def test_as_ilu0_factors_as_row_sum_dilu():
    # DILU_DInv_crs tests ILU_method == ILUD only: ILU(0) takes the row-sum branch
    rp, col, val = poisson2d(16)
    n = rp.shape[0] - 1
    r = np.random.default_rng(5).standard_normal(n)
    a = orc.AsPrecon(rp, col, val, ilu_method=0, nblocks=3, addL=7, iflagAS=1).apply(r)
    b = orc.AsPrecon(rp, col, val, ilu_method=2, nblocks=3, addL=7, iflagAS=1).apply(r)
    assert np.array_equal(a, b)


def test_as_basic_is_symmetric_and_helps_cg():
    # BASIC with ILUD local factors (D+L) D^-1 (D+L^T) of an SPD matrix is a
    # symmetric operator; PCG with it needs fewer iterations than block Jacobi
    rp, col, val = poisson2d(32)
    n = rp.shape[0] - 1
    M = orc.AsPrecon(rp, col, val, ilu_method=1, nblocks=8, addL=32, iflagAS=1)
    E = np.stack([M.apply(e) for e in np.eye(n)[:200]])
    assert np.allclose(E[:, :200], E[:, :200].T, rtol=1e-12, atol=1e-14)
    b = np.random.default_rng(6).standard_normal(n)
    _, ha, na, ia = orc.pcg_factor(rp, col, val, b, M, itrmax=2000, rtol=1e-10)
    _, hb, nb_, ib = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=1, nblocks=8, itrmax=2000, rtol=1e-10)
    assert ia == 0 and ib == 0 and na < nb_
