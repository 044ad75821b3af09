"""CPU tests of the oracle's restarted GMRES(m) (src/pgmres.f90:62-525).

The reference ships no GMRES golden output, so the oracle is pinned by what
the mathematics fixes: within one restart cycle, the residual after j steps
of right-preconditioned GMRES is the minimum of ||r_c - A M^-1 z|| over the
Krylov space K_j(A M^-1, r_c); it is computed here independently with a
re-orthogonalised Arnoldi basis and LAPACK least squares.  The solution
update (back substitution, basis combination, preconditioning) is pinned by
the true residual of the returned x, and the loop bookkeeping (niter, iret,
where the history is written) by the DO-loop semantics of the reference,
worked out by hand in test_gmres_exit_semantics."""
import ctypes

import numpy as np
import pytest
import scipy.linalg

import oracle as orc
import parcel_hip as ph
from problems import convdiff3d


def min_residuals(A, minv, r0, m):
    """||r0||, then min over z in K_j(A M^-1, r0) of ||r0 - A M^-1 z||, j=1..m."""
    n = r0.shape[0]
    beta = np.linalg.norm(r0)
    V = np.zeros((n, m + 1))
    V[:, 0] = r0 / beta
    AZ = np.zeros((n, m))
    out = [beta]
    for j in range(m):
        w = A @ minv(np.ascontiguousarray(V[:, j]))
        AZ[:, j] = w
        for _ in range(2):  # classical Gram-Schmidt, twice
            w = w - V[:, : j + 1] @ (V[:, : j + 1].T @ w)
        V[:, j + 1] = w / np.linalg.norm(w)
        # least squares in the image of the orthonormal Krylov basis
        Q, _ = np.linalg.qr(AZ[:, : j + 1])
        res = r0 - Q @ (Q.T @ r0)
        out.append(np.linalg.norm(res))
    return np.array(out)


def _precon(rp, col, val, ipreflag, **kw):
    if ipreflag == 0:
        return lambda v: v.copy()
    P = orc.Precon(rp, col, val, ipreflag, **kw)
    return P.apply


CASES = [
    ("none", dict(ipreflag=0)),
    ("pj", dict(ipreflag=1)),
    ("bj_ilu0", dict(ipreflag=2, ilu_method=0, nblocks=4)),
    ("bj_dilu", dict(ipreflag=2, ilu_method=1, nblocks=3)),
]


@pytest.mark.parametrize("gs", [1, 2])
@pytest.mark.parametrize("name,kw", CASES)
def test_gmres_first_cycle_is_minimal_residual(name, kw, gs):
    rp, col, val, A = convdiff3d(6, peclet=2.0)
    n = A.shape[0]
    b = ph.uniform(n, 11)
    m = 8
    # itrmax = m + 1: the first cycle runs all m steps (pgmres.f90:410-416)
    x, hist, niter, iret = orc.gmres(rp, col, val, b, itrmax=m + 1, rtol=1e-30, m=m, gsflag=gs, **kw)
    assert niter == m + 1 and iret == 0
    pk = {k: v for k, v in kw.items() if k != "ipreflag"}
    ref = min_residuals(A, _precon(rp, col, val, kw["ipreflag"], **pk), b, m) / np.linalg.norm(b)
    np.testing.assert_allclose(hist[: m + 1], ref, rtol=1e-9, atol=0)
    # the update: the true residual of x equals the last Givens residual
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert abs(true - hist[m]) <= 1e-9 * hist[m]


@pytest.mark.parametrize("gs", [1, 2])
def test_gmres_restarts_minimise_from_each_cycle_start(gs):
    rp, col, val, A = convdiff3d(5, peclet=3.0)
    n = A.shape[0]
    b = ph.uniform(n, 12)
    m = 3
    kw = dict(ipreflag=2, ilu_method=1, nblocks=2)
    minv = _precon(rp, col, val, 2, ilu_method=1, nblocks=2)
    bnorm = np.linalg.norm(b)
    x_prev = np.zeros(n)
    for c in range(1, 4):
        # itrmax = c*m + 1 stops after exactly c full cycles
        x_c, hist, niter, iret = orc.gmres(rp, col, val, b, itrmax=c * m + 1, rtol=1e-30, m=m, gsflag=gs, **kw)
        assert niter == c * m + 1 and iret == 0
        ref = min_residuals(A, minv, b - A @ x_prev, m) / bnorm
        np.testing.assert_allclose(hist[(c - 1) * m: c * m + 1], ref, rtol=1e-8, atol=0)
        true = np.linalg.norm(b - A @ x_c) / bnorm
        assert abs(true - hist[c * m]) <= 1e-8 * hist[c * m]
        x_prev = x_c


def test_gmres_mgs_and_cgs_differ_only_by_rounding():
    rp, col, val, A = convdiff3d(6, peclet=1.0)
    b = ph.uniform(A.shape[0], 13)
    _, h1, n1, _ = orc.gmres(rp, col, val, b, itrmax=31, rtol=1e-30, m=10, gsflag=1, ipreflag=1)
    _, h2, n2, _ = orc.gmres(rp, col, val, b, itrmax=31, rtol=1e-30, m=10, gsflag=2, ipreflag=1)
    assert n1 == n2 == 31
    np.testing.assert_allclose(h1, h2, rtol=1e-7)
    assert not np.array_equal(h1, h2)
    # any GS flag other than MGS takes the CGS branch (pgmres.f90:336)
    _, h5, _, _ = orc.gmres(rp, col, val, b, itrmax=31, rtol=1e-30, m=10, gsflag=5, ipreflag=1)
    assert np.array_equal(h2, h5)


def test_gmres_converges_to_solution():
    rp, col, val, A = convdiff3d(8, peclet=1.0)
    b = ph.uniform(A.shape[0], 14)
    x, hist, niter, iret = orc.gmres(rp, col, val, b, itrmax=2000, rtol=1e-10, m=20, gsflag=1, ipreflag=2,
                                     ilu_method=0, nblocks=4)
    assert iret == 0 and niter < 2000
    assert hist[-1] < 1e-10
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9
    xs = scipy.linalg.solve(A.toarray(), b)
    assert np.allclose(x, xs, rtol=0, atol=1e-8 * np.abs(xs).max())


@pytest.mark.parametrize("itrmax,m,niter,iret,written", [
    # itrmax < m+1: the first cycle never meets itr+inner == itrmax, runs m
    # steps and the DO loop ends with itr = 1+m >= itrmax
    (1, 4, 1, 1, 5),
    # itr + inner == itrmax inside the first cycle: inner exits at 3; iret is
    # 0 because it only tests itrmax <= itr (pgmres.f90:501)
    (4, 4, 4, 0, 4),
    # the first cycle completes (inner = m+1 = 5), itr + inner = 6 == itrmax
    # ends the outer loop after the update; reshistory(6) is never written
    (6, 4, 6, 0, 5),
    (9, 4, 9, 0, 9),
    (0, 4, 0, 1, 0),
])
def test_gmres_exit_semantics(itrmax, m, niter, iret, written):
    rp, col, val, A = convdiff3d(4, peclet=1.0)
    b = ph.uniform(A.shape[0], 15)
    n = A.shape[0]
    hist = np.full(max(itrmax, m + 1, 1), -1.0)
    it, ret = ctypes.c_int(0), ctypes.c_int(0)
    x = np.zeros(n)
    rc = orc.load().oracle_gmres(n, orc._p(rp), orc._p(col), orc._p(val), orc._p(b), orc._p(x), 0, 1, 1, 1,
                                 itrmax, 1e-30, m, 1, ctypes.byref(orc.TILED), orc._p(hist), ctypes.byref(it),
                                 ctypes.byref(ret))
    assert rc == 0
    assert (it.value, ret.value) == (niter, iret)
    assert int(np.sum(hist != -1.0)) == written
    if written:
        assert hist[0] == 1.0  # x0 = 0: r0 = b


def test_gmres_converged_at_restart_check():
    # an exact initial guess: the restart test exits before any step,
    # niter = min(itrmax, itr + inner) = 2 with inner still 1
    rp, col, val, A = convdiff3d(4)
    xs = ph.uniform(A.shape[0], 16)
    b = A @ xs
    x, hist, niter, iret = orc.gmres(rp, col, val, b, x0=xs, itrmax=50, rtol=1e-8, m=5)
    assert niter == 2 and iret == 0
    assert np.array_equal(x, xs)


def test_gmres_planted_mistake_is_caught():
    """Left instead of right preconditioning (a plausible slip) changes the
    minimised quantity; the first-cycle check above would catch it."""
    rp, col, val, A = convdiff3d(6, peclet=2.0)
    b = ph.uniform(A.shape[0], 11)
    m = 6
    minv = _precon(rp, col, val, 2, ilu_method=0, nblocks=4)
    right = min_residuals(A, minv, b, m)
    # left preconditioning minimises ||M^-1 r||; its true residuals differ
    n = A.shape[0]
    MA = np.column_stack([minv(np.ascontiguousarray(A[:, [i]].toarray().ravel())) for i in range(n)])
    r0 = minv(b)
    left = [np.linalg.norm(b)]
    for j in range(1, m + 1):
        K = np.column_stack([np.linalg.matrix_power(MA, k) @ r0 for k in range(j)])
        z = np.linalg.lstsq(MA @ K, r0, rcond=None)[0]
        left.append(np.linalg.norm(b - A @ (K @ z)))
    _, hist, _, _ = orc.gmres(rp, col, val, b, itrmax=m + 1, rtol=1e-30, m=m, ipreflag=2, ilu_method=0, nblocks=4)
    bn = np.linalg.norm(b)
    assert np.max(np.abs(hist[1:] - right[1:] / bn) / (right[1:] / bn)) < 1e-9
    assert np.max(np.abs(np.array(left[1:]) - right[1:]) / right[1:]) > 1e-3
