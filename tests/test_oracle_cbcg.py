"""CPU tests of the oracle's Chebyshev-basis CG (KSP_pCBCG,
src/pcbcg.f90:1182-2025), its power-method bound
(src/power_method.f90:188-437) and the symmetric pseudo-inverse standing in
for moore_penrose (src/Moore_Penrose.f90:3-67).

No reference fixture exists for solver 5, so the oracle is pinned by what the
mathematics fixes:
* k-step CBCG is an s-step CG: every outer step minimises the A-norm error
  over the same Krylov space as k steps of PCG, so its history entry j must
  equal PCG's residual after k*j iterations (any basis of the space, hence any
  eigenvalue bounds, gives the same iterate), checked before rounding
  separates them;
* the pseudo-inverse against numpy's SVD-based pinv, on definite, indefinite
  and exactly singular symmetric matrices;
* the power method against numpy's eigenvalues of A M^-1, from a given start
  and from x0 = 0 (the all-ones start);
* the loop bookkeeping against the DO-loop semantics of :1560-1980
  (do itr = 1, itrmax/k; niter = min(itrmax, itr); iret = itrmax <= itr);
* the full power method of parcel_dpower (Emax, the eigenpair error
  ||A x - Emax x|| / ||Emax x|| and the eigenvector) against its closed form
  on diagonal operators, where K steps from x0 give x = A^K x0 / ||A^K x0||
  and Emax = sum d^(2K-1) / sum d^(2K-2) (exact rationals), and against the
  closed-form top eigenpair of the reference's 3D Poisson matrix,
  lambda = -6 - 6 cos(pi/(N+1)), whose value for N = 160 the reference's own
  eigen run prints (src/example/template_Eigen_cpu/STDOUT: -11.9988577649884).
"""
import math
from fractions import Fraction
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import parcel_hip as ph
from problems import convdiff3d


def _dense(rp, col, val):
    n = rp.shape[0] - 1
    return sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n)).toarray()


def _minv_dense(rp, col, val, ipreflag, nblocks):
    n = rp.shape[0] - 1
    pre = orc.Precon(rp, col, val, ipreflag, 0, nblocks)
    return np.column_stack([pre.apply(np.ascontiguousarray(e)) for e in np.eye(n)])


def test_sym_pinv_matches_numpy():
    rng = np.random.default_rng(5)
    for k in (1, 2, 3, 4, 6, 8):
        V = rng.standard_normal((30, k))
        spd = V.T @ V
        assert np.allclose(orc.sym_pinv(spd), np.linalg.pinv(spd), rtol=1e-11, atol=1e-12)
        B = rng.standard_normal((k, k))
        ind = B + B.T
        assert np.allclose(orc.sym_pinv(ind), np.linalg.pinv(ind), rtol=1e-9, atol=1e-10)
    # exactly singular (a zero block that no rotation touches): the zero
    # eigenvalues are not inverted, as DGESVD's sigma > 0 test leaves them
    S3 = spd[:3, :3]
    A = np.zeros((5, 5))
    A[:3, :3] = S3
    P = orc.sym_pinv(A)
    assert np.allclose(P[:3, :3], np.linalg.inv(S3), rtol=1e-11) and not P[3:, :].any() and not P[:, 3:].any()
    # only the upper triangle is read
    L = np.tril(rng.standard_normal((5, 5)), -1) * 1e3
    S = spd[:5, :5] if spd.shape[0] >= 5 else None
    if S is not None:
        assert np.array_equal(orc.sym_pinv(np.triu(S) + L), orc.sym_pinv(S))
    # diagonal matrix with a zero: the zero is not inverted
    D = np.diag([2.0, 0.0, -4.0])
    assert np.array_equal(orc.sym_pinv(D), np.diag([0.5, 0.0, -0.25]))


def test_power_method_converges_to_largest_eigenvalue():
    # diagonal operator with a separated top eigenvalue: the Rayleigh quotient
    # error after 100 steps is (lambda2/lambda1)^200 relative
    n = 40
    d = np.linspace(1.0, 5.0, n)
    d[-1] = 10.0
    rp = np.arange(1, n + 2, dtype=np.int32)
    col = np.arange(1, n + 1, dtype=np.int32)
    e = orc.power_emax(rp, col, d, np.zeros(n), ipreflag=0, itrmax=100)
    assert abs(e - 10.0) < 1e-12 * 10
    # SPD Poisson with point Jacobi: eigenvalues of A M^-1 are real
    rp, col, val = ph.poisson3d_spd(6, 5, 4)
    Ad = _dense(rp, col, val)
    lam = np.linalg.eigvals(Ad @ _minv_dense(rp, col, val, 1, 1))
    top = np.max(np.abs(lam))
    rng = np.random.default_rng(3)
    e0 = orc.power_emax(rp, col, val, rng.standard_normal(rp.shape[0] - 1), ipreflag=1, itrmax=100)
    e1 = orc.power_emax(rp, col, val, np.zeros(rp.shape[0] - 1), ipreflag=1, itrmax=100)
    assert abs(e0 - top) < 2e-2 * top
    # the all-ones start has no component on the top mode of this grid (the
    # highest modes in x and z are odd functions), so it settles lower
    assert 0 < e1 < top
    assert e0 <= top * (1 + 1e-12)  # Rayleigh quotient of a symmetric pencil


@pytest.mark.parametrize("k", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("pre", [(0, 1), (1, 1), (2, 4)])
def test_cbcg_is_s_step_cg(k, pre):
    ipreflag, nb = pre
    rp, col, val = ph.poisson3d_spd(7, 6, 5)
    n = rp.shape[0] - 1
    rng = np.random.default_rng(11)
    b = rng.standard_normal(n)
    Ad = _dense(rp, col, val)
    lam = np.real(np.linalg.eigvals(Ad @ _minv_dense(rp, col, val, ipreflag, nb)))
    _, hc, _, _ = orc.pcg(rp, col, val, b, ipreflag=ipreflag, ilu_method=0, nblocks=nb, itrmax=400, rtol=1e-14)
    for bounds in ((lam.max(), lam.min()), (lam.max() * 1.1, 0.0)):
        x, hb, nit, ret = orc.cbcg_bounds(rp, col, val, b, ipreflag=ipreflag, ilu_method=0, nblocks=nb,
                                          itrmax=400, rtol=1e-9, k=k, emax=bounds[0], emin=bounds[1])
        assert hb[0] == pytest.approx(1.0)
        # PCG's history entry i is ||r_i||/||b|| (entry 0 the start); CBCG's
        # entry j is ||r_{kj}||/||b||
        m = min(len(hb) - 1, (len(hc) - 1) // k)
        upto = [j for j in range(1, m + 1) if hc[k * j] > 1e-7]
        assert len(upto) >= 2
        for j in upto:
            assert hb[j] == pytest.approx(hc[k * j], rel=1e-5), (k, bounds, j)
        assert ret == 0 and hb[-1] < 1e-9
        r = b - sp.csr_matrix(Ad) @ x
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8


def test_cbcg_power_bound_solves_nonsymmetric_problem():
    rp, col, val, _ = convdiff3d(6, peclet=0.3)
    n = rp.shape[0] - 1
    b = np.random.default_rng(2).standard_normal(n)
    x, h, nit, ret, (emax, emin) = orc.cbcg(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=2, itrmax=600,
                                           rtol=1e-8, k=2, eigflag=1, pm_itrmax=50)
    assert emin == 0.0 and emax > 0
    assert ret == 0 and h[-1] < 1e-8
    Ad = _dense(rp, col, val)
    assert np.linalg.norm(b - Ad @ x) / np.linalg.norm(b) < 1e-7


def test_cbcg_exit_semantics():
    rp, col, val = ph.poisson3d_spd(8, 8, 8)
    n = rp.shape[0] - 1
    b = np.ones(n)
    # itrmax = 8, k = 4: two outer steps, loop ends at itr = 3
    x, h, nit, ret, _ = orc.cbcg(rp, col, val, b, ipreflag=1, itrmax=8, rtol=1e-30, k=4, eigflag=1, pm_itrmax=5)
    assert nit == 3 and ret == 0 and h.shape == (3,) and h[2] == 0.0 and h[0] == 1.0 and 0 < h[1] < 1
    # k = 1: itrmax outer steps, niter = itrmax, iret = 1
    x, h, nit, ret, _ = orc.cbcg(rp, col, val, b, ipreflag=1, itrmax=5, rtol=1e-30, k=1, eigflag=1, pm_itrmax=5)
    assert nit == 5 and ret == 1 and np.all(h > 0)
    # unknown eigflag: Emax = 1, Emin = 0 (still a valid Krylov basis)
    x, h, nit, ret, (emax, emin) = orc.cbcg(rp, col, val, b, ipreflag=1, itrmax=400, rtol=1e-8, k=3, eigflag=0)
    assert (emax, emin) == (0.0, 0.0) and ret == 0 and h[-1] < 1e-8
    # converged at the start: x0 = exact solution, history (0) in step 1
    xe = np.linalg.solve(_dense(rp, col, val), b)
    x, h, nit, ret, _ = orc.cbcg(rp, col, val, b, x0=xe, ipreflag=1, itrmax=40, rtol=1e-6, k=4, eigflag=1,
                                 pm_itrmax=3)
    assert nit == 1 and ret == 0 and h[0] < 1e-6


def test_cbcg_caarnoldi_flag_is_the_power_method():
    # KSP_pCAARNOLDI's GPU-build body is power_method_symMatrix
    # (src/pcagmres.f90:2647-2651): flag 2 gives flag 1's bound and history
    rp, col, val = ph.poisson3d_spd(8, 7, 6)
    b = np.random.default_rng(9).standard_normal(rp.shape[0] - 1)
    r1 = orc.cbcg(rp, col, val, b, ipreflag=1, itrmax=200, rtol=1e-9, k=4, eigflag=1, pm_itrmax=20)
    r2 = orc.cbcg(rp, col, val, b, ipreflag=1, itrmax=200, rtol=1e-9, k=4, eigflag=2, pm_itrmax=20)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1]) and r1[2:] == r2[2:]



def _power_closed_form(d, x0, K):
    """K power steps on diag(d) from x0 in exact rationals: (Emax, err, x)."""
    d = [Fraction(v) for v in d]
    x0 = [Fraction(v) for v in x0]
    yk1 = [di ** (K - 1) * xi for di, xi in zip(d, x0)]  # A^(K-1) x0
    yk = [di ** K * xi for di, xi in zip(d, x0)]  # A^K x0
    emax = sum(a * di * a for a, di in zip(yk1, d)) / sum(a * a for a in yk1)
    nk = sum(a * a for a in yk)
    num = sum((di - emax) ** 2 * a * a for di, a in zip(d, yk))
    err = math.sqrt(num / nk) / math.sqrt(emax * emax)
    x = np.array([float(a) for a in yk]) / math.sqrt(nk)
    return float(emax), err, x


@pytest.mark.parametrize("d,K", [((1.0, 2.0, 3.0), 5), ((-3.0, 1.0, 2.0, 2.5), 7), ((4.0, -1.0), 1)])
def test_power_eigen_closed_form_on_diagonal_operators(d, K):
    n = len(d)
    rp = np.arange(1, n + 2, dtype=np.int32)
    col = np.arange(1, n + 1, dtype=np.int32)
    val = np.array(d)
    for x0 in (np.array([1.0, 0.5, -2.0, 3.0][:n]), np.zeros(n)):
        start = x0 if x0.any() else np.ones(n)  # a zero start is replaced by all ones
        e, err, vec = orc.power_eigen(rp, col, val, x0, ipreflag=0, itrmax=K)
        ce, cerr, cvec = _power_closed_form(d, start, K)
        assert abs(e - ce) <= 1e-14 * abs(ce)
        assert abs(err - cerr) <= 1e-12 * cerr
        assert np.allclose(vec, cvec, rtol=1e-14, atol=1e-15)
    # the error is relative to ||Emax x||, not ||A x||: the two differ before convergence
    e, err, vec = orc.power_eigen(rp, col, val, np.array([1.0, 0.5, -2.0, 3.0][:n]), itrmax=K)
    ax = np.linalg.norm(val * vec)
    assert abs(ax - abs(e)) > 1e-6 * ax


def test_power_eigen_poisson_top_eigenpair():
    # the closed form printed by the reference's eigen run on its 160^3 Poisson
    assert abs(-6.0 - 6.0 * math.cos(math.pi / 161) - (-11.9988577649884)) < 1e-12
    N = 6
    rp, col, val = ph.poisson3d_parcel(N, N, N)
    lam = -6.0 - 6.0 * math.cos(math.pi / (N + 1))
    rng = np.random.default_rng(11)
    e, err, vec = orc.power_eigen(rp, col, val, rng.standard_normal(N ** 3), itrmax=2000)
    assert abs(e - lam) < 1e-12 * abs(lam)
    assert err < 1e-12
    # the top mode: sin(k pi i / (N+1)) with k = N in each direction
    s = np.sin(N * np.pi * np.arange(1, N + 1) / (N + 1))
    v = np.einsum("i,j,k->kji", s, s, s).ravel()
    v /= np.linalg.norm(v)
    assert abs(abs(vec @ v) - 1.0) < 1e-12 and abs(np.linalg.norm(vec) - 1.0) < 1e-14


def test_power_eigen_error_with_preconditioner_uses_a():
    # with M the iteration is on A M^-1, the error check on A alone
    # (calc_Matvec after the loop, src/power_method.f90:366-398)
    rp, col, val = ph.poisson3d_spd(5, 4, 3)
    n = rp.shape[0] - 1
    rng = np.random.default_rng(4)
    x0 = rng.standard_normal(n)
    e, err, vec = orc.power_eigen(rp, col, val, x0, ipreflag=1, itrmax=40)
    Ad = _dense(rp, col, val)
    assert abs(e - orc.power_emax(rp, col, val, x0, ipreflag=1, itrmax=40)) == 0.0
    assert abs(np.linalg.norm(vec) - 1.0) < 1e-14
    ref = np.linalg.norm(Ad @ vec - e * vec) / np.linalg.norm(e * vec)
    assert abs(err - ref) < 1e-10 * ref
