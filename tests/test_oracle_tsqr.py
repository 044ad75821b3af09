"""CPU tests of the oracle's TSQR R factor (TSQR_Ronly_,
src/QRDecomposition.f90:887-1002; src/TSQR.f90:114-498,836-927) and of
CA-GMRES with QRflag = 3.

Pins (no reference fixture holds a TSQR factor):
* the R factor of a full-rank V with a positive diagonal is unique, so the
  oracle's R must equal numpy's (LAPACK) QR factor with its rows' signs made
  positive, to rounding, for widths 1..16 and lengths that leave a ragged last
  panel, one panel per wavefront and many panels per wavefront;
* Householder stability: on a V with condition number 1e7, Q = V R^-1 from
  TSQR is orthogonal to ~1e-8 while CholeskyQR's loses orthogonality
  (squared condition number);
* House_local_Ronly_omp (the second level) against numpy up to row signs;
* CA-GMRES with TSQR reaches the GMRES minimal residual, and falls back to
  CholeskyQR2 (bit for bit) when n <= 64, as QRDecomposition_ does.
"""
import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph
from problems import convdiff3d
from test_oracle_gmres import min_residuals, _precon


def _np_r(V):
    R = np.linalg.qr(V, mode="r")
    d = np.sign(np.diag(R))
    d[d == 0] = 1.0
    return d[:, None] * R


@pytest.mark.parametrize("n,nc", [(65, 1), (100, 3), (1000, 5), (4097, 9), (70000, 16), (200003, 4)])
def test_tsqr_matches_lapack_r(n, nc):
    rng = np.random.default_rng(n + nc)
    V = rng.standard_normal((n, nc))
    R = orc.tsqr_r(V)
    assert np.all(np.diag(R) > 0)
    assert np.allclose(np.tril(R, -1), 0.0)
    Rn = _np_r(V)
    assert np.abs(R - Rn).max() <= 1e-12 * np.abs(Rn).max()


def test_tsqr_wave_count():
    assert orc.load().oracle_tsqr_waves(64) == 1
    assert orc.load().oracle_tsqr_waves(65) == 2
    assert orc.load().oracle_tsqr_waves(64 * 1024) == 1024
    assert orc.load().oracle_tsqr_waves(64 * 5000) == 1024


def test_tsqr_is_backward_stable_where_cholqr_is_not():
    rng = np.random.default_rng(7)
    n, nc = 20000, 6
    U, _ = np.linalg.qr(rng.standard_normal((n, nc)))
    W, _ = np.linalg.qr(rng.standard_normal((nc, nc)))
    V = U @ np.diag(np.logspace(0, -7, nc)) @ W
    R = orc.tsqr_r(V)
    Q = V @ np.linalg.inv(R)
    e_tsqr = np.abs(Q.T @ Q - np.eye(nc)).max()
    Rc = np.linalg.cholesky(V.T @ V).T
    Qc = V @ np.linalg.inv(Rc)
    e_chol = np.abs(Qc.T @ Qc - np.eye(nc)).max()
    assert e_tsqr < 1e-8
    assert e_chol > 1e3 * e_tsqr


def test_house_r_matches_lapack_up_to_row_signs():
    rng = np.random.default_rng(3)
    for m, s in ((5, 5), (40, 3), (300, 8)):
        A = rng.standard_normal((m, s))
        R = orc.house_r(A)
        Rn = np.linalg.qr(A, mode="r")
        d = np.sign(np.diag(R)) * np.sign(np.diag(Rn))
        assert np.allclose(R, d[:, None] * Rn, rtol=1e-12, atol=1e-12 * np.abs(Rn).max())


def test_tsqr_zero_column_stays_finite():
    rng = np.random.default_rng(4)
    V = rng.standard_normal((500, 4))
    V[:, 2] = 0.0
    R = orc.tsqr_r(V)
    assert np.all(np.isfinite(R))
    assert np.allclose(R[:, 2], 0.0)


@pytest.mark.parametrize("s,t", [(4, 1), (3, 3)])
def test_cagmres_tsqr_first_cycle_is_minimal_residual(s, t):
    rp, col, val, A = convdiff3d(6, peclet=2.0)
    b = ph.uniform(A.shape[0], 35)
    m = s * t
    x, hist, niter, iret = orc.cagmres(rp, col, val, b, itrmax=m + 1, rtol=1e-30, s=s, t=t, basis=0, qrflag=3,
                                       ipreflag=1)
    ref = min_residuals(A, _precon(rp, col, val, 1), b, m) / np.linalg.norm(b)
    np.testing.assert_allclose(hist, ref, rtol=1e-6, atol=0)
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert abs(true - hist[-1]) <= 1e-6 * hist[-1]


def test_cagmres_tsqr_falls_back_to_cholqr2_for_short_vectors():
    rp, col, val, A = convdiff3d(4, peclet=1.0)  # n = 64
    b = ph.uniform(A.shape[0], 36)
    r3 = orc.cagmres(rp, col, val, b, itrmax=40, rtol=1e-12, s=4, t=2, basis=1, qrflag=3, ipreflag=1)
    r5 = orc.cagmres(rp, col, val, b, itrmax=40, rtol=1e-12, s=4, t=2, basis=1, qrflag=5, ipreflag=1)
    assert np.array_equal(r3[0], r5[0]) and np.array_equal(r3[1], r5[1]) and r3[2:] == r5[2:]
