"""CPU tests of the oracle's CA-GMRES(s,t) (src/pcagmres.f90:1557-2517) and
of the small dense kernels it uses.

No reference fixture exists for solver 4, so:
* the dense kernels are checked against numpy (LAPACK) and closed forms;
* the Leja permutation is checked against a brute-force greedy in exact
  rational arithmetic that follows the reference's rule
  (src/LejaOrdering.f90:111-277: stable sort by modulus, the second point is
  the next in sorted order because the candidate product runs over the first
  L-1 chosen points only, conjugate pairs taken together);
* every restart cycle's residual history must equal the minimal residual
  over the right-preconditioned Krylov space from that cycle's start (any
  basis of the space -- monomial or Newton -- gives the same minimum), and
  the returned x must reproduce the last recorded residual.
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph
from problems import convdiff3d
from test_oracle_gmres import min_residuals, _precon


# --- dense kernels -----------------------------------------------------------

def test_gemm_trmm_match_numpy():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((7, 5))
    B = rng.standard_normal((5, 4))
    assert np.allclose(orc.gemm_nn(A, B), A @ B, rtol=1e-14, atol=1e-14)
    U = np.triu(rng.standard_normal((6, 6)))
    X = rng.standard_normal((6, 6))
    assert np.allclose(orc.trmm_lunn(U, X), U @ X, rtol=1e-13, atol=1e-13)


def test_cholesky_and_triangular_inverse():
    rng = np.random.default_rng(2)
    for n in (1, 2, 5, 9, 17):
        V = rng.standard_normal((40, n))
        G = V.T @ V
        R, info = orc.chol_upper(np.triu(G))
        assert info == 0
        assert np.allclose(R.T @ R, G, rtol=1e-12, atol=1e-12 * np.abs(G).max())
        assert np.all(np.diag(R) > 0)
        Ri = orc.trinv_upper(R)
        assert np.allclose(Ri @ R, np.eye(n), atol=1e-10)
    # a non-positive pivot is reported at its 1-based column, as DPOTF2 does
    G = np.array([[4.0, 2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    _, info = orc.chol_upper(np.triu(G))
    assert info == 2


def _hessenberg(n, rng):
    return np.triu(rng.standard_normal((n, n)), -1)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 12, 30])
def test_hessenberg_eigenvalues_match_numpy(n):
    rng = np.random.default_rng(n)
    for _ in range(10):
        H = _hessenberg(n, rng)
        wr, wi = orc.hqr_eigs(H)
        got = np.sort_complex(wr + 1j * wi)
        ref = np.sort_complex(np.linalg.eigvals(H))
        assert np.max(np.abs(got - ref)) <= 1e-11 * max(1.0, np.abs(ref).max())
        for i in range(n):  # conjugate pairs adjacent, positive imaginary part first
            if wi[i] > 0:
                assert wi[i + 1] == -wi[i] and wr[i + 1] == wr[i]


def test_hessenberg_eigenvalues_closed_forms():
    # companion matrix (upper Hessenberg after transposition) of prod (z - k)
    roots = np.arange(1.0, 9.0)
    c = np.poly(roots)
    C = np.zeros((8, 8))
    C[0, :] = -c[1:]
    C[1:, :-1] = np.eye(7)
    wr, wi = orc.hqr_eigs(C)
    assert np.allclose(np.sort(wr), roots, rtol=1e-8) and np.all(wi == 0)
    # rotation-scaling block: eigenvalues 2 +- 3i
    wr, wi = orc.hqr_eigs(np.array([[2.0, -3.0], [3.0, 2.0]]))
    assert np.allclose(wr, [2, 2]) and np.allclose(wi, [3, -3])


# --- Leja ordering -------------------------------------------------------------

def leja_brute_force(re, im):
    """The reference's modified Leja greedy, in exact rational arithmetic;
    returns the chosen indices into the modulus-sorted list."""
    z = [(Fraction(a), Fraction(b)) for a, b in zip(re, im)]
    # stable descending sort by modulus (insertion sort moves an element left
    # only past strictly smaller moduli)
    order = list(range(len(z)))
    mod2 = lambda p: p[0] * p[0] + p[1] * p[1]
    srt = []
    for i in order:
        k = len(srt)
        while k > 0 and mod2(z[srt[k - 1]]) < mod2(z[i]):
            k -= 1
        srt.insert(k, i)
    zs = [z[i] for i in srt]
    n = len(zs)
    if zs[0][1] > 0:
        out = [0, 1]
    else:
        out = [0]
    while len(out) < n:
        L = len(out)
        best, k = Fraction(0), 0
        first = True
        for j in range(n):
            if j in out:
                score = Fraction(0)
            else:
                score = Fraction(1)
                for q in range(L - 1):
                    d = (zs[j][0] - zs[out[q]][0], zs[j][1] - zs[out[q]][1])
                    score *= mod2(d)
            if first:
                best, k, first = score, j, False
            elif score > best:
                best, k = score, j
        if zs[k][1] != 0:
            out += [k - 1, k] if zs[k][1] < 0 else [k, k + 1]
        else:
            out.append(k)
    return [srt[i] for i in out]


@pytest.mark.parametrize("seed", range(6))
def test_leja_permutation_real(seed):
    rng = np.random.default_rng(seed)
    re = np.round(rng.uniform(-10, 10, size=8), 3)
    im = np.zeros(8)
    perm = leja_brute_force(re, im)
    ore, oim = orc.leja(re, im)
    np.testing.assert_allclose(ore, re[perm], rtol=1e-12, atol=0)
    assert np.all(oim == 0)


@pytest.mark.parametrize("seed", range(6))
def test_leja_permutation_complex_pairs(seed):
    rng = np.random.default_rng(100 + seed)
    vals = []
    for _ in range(3):  # conjugate pairs, positive imaginary part first
        a, b = np.round(rng.uniform(-5, 5), 3), np.round(rng.uniform(0.5, 5), 3)
        vals += [(a, b), (a, -b)]
    for _ in range(2):
        vals.append((np.round(rng.uniform(-5, 5), 3), 0.0))
    re = np.array([v[0] for v in vals])
    im = np.array([v[1] for v in vals])
    perm = leja_brute_force(re, im)
    ore, oim = orc.leja(re, im)
    np.testing.assert_allclose(ore, re[perm], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oim, im[perm], rtol=1e-12, atol=1e-12)
    for i in range(len(vals)):  # pairs stay adjacent, + before -
        if oim[i] > 0:
            assert oim[i + 1] < 0


def test_leja_second_point_is_next_in_sorted_order():
    # mleja_prod multiplies over the first L-1 points: at L = 1 every
    # candidate scores 1 and the first unchosen (sorted) index wins
    re, im = orc.leja([1.0, 5.0, -4.9, 0.1], [0.0] * 4)
    np.testing.assert_allclose(re[:2], [5.0, -4.9], rtol=1e-14)


# --- CA-GMRES -----------------------------------------------------------------

@pytest.mark.parametrize("qr", [1, 2, 4, 5])
@pytest.mark.parametrize("pre", [dict(ipreflag=0), dict(ipreflag=2, ilu_method=0, nblocks=4)])
def test_cagmres_first_cycle_is_minimal_residual(qr, pre):
    rp, col, val, A = convdiff3d(6, peclet=2.0)
    b = ph.uniform(A.shape[0], 21)
    s = 5
    # itrmax = s + 1: the first cycle runs all s steps and ends the solve
    x, hist, niter, iret = orc.cagmres(rp, col, val, b, itrmax=s + 1, rtol=1e-30, s=s, qrflag=qr, **pre)
    assert niter == s + 1 and iret == 0
    pk = {k: v for k, v in pre.items() if k != "ipreflag"}
    ref = min_residuals(A, _precon(rp, col, val, pre["ipreflag"], **pk), b, s) / np.linalg.norm(b)
    np.testing.assert_allclose(hist, ref, rtol=1e-7, atol=0)
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert abs(true - hist[-1]) <= 1e-7 * hist[-1]


@pytest.mark.parametrize("basis,qr", [(0, 4), (1, 4), (1, 5), (1, 1), (1, 2)])
def test_cagmres_restarts_minimise_from_each_cycle_start(basis, qr):
    rp, col, val, A = convdiff3d(5, peclet=3.0)
    n = A.shape[0]
    b = ph.uniform(n, 22)
    s = 4
    kw = dict(ipreflag=2, ilu_method=1, nblocks=2)
    minv = _precon(rp, col, val, 2, ilu_method=1, nblocks=2)
    bnorm = np.linalg.norm(b)
    x_prev = np.zeros(n)
    for c in range(1, 4):
        x_c, hist, niter, iret = orc.cagmres(rp, col, val, b, itrmax=c * s + 1, rtol=1e-30, s=s, basis=basis,
                                             qrflag=qr, **kw)
        assert niter == c * s + 1 and iret == 0
        ref = min_residuals(A, minv, b - A @ x_prev, s) / bnorm
        np.testing.assert_allclose(hist[(c - 1) * s: c * s + 1], ref, rtol=1e-7, atol=0)
        true = np.linalg.norm(b - A @ x_c) / bnorm
        assert abs(true - hist[c * s]) <= 1e-7 * hist[c * s]
        x_prev = x_c


def test_cagmres_newton_basis_converges_where_monomial_stalls():
    # large s: the monomial basis loses rank in CholeskyQR, the Newton basis
    # with Leja-ordered Ritz shifts keeps converging
    rp, col, val, A = convdiff3d(8, peclet=1.0)
    b = ph.uniform(A.shape[0], 23)
    x, hist, niter, iret = orc.cagmres(rp, col, val, b, itrmax=3000, rtol=1e-10, s=12, basis=1, qrflag=5,
                                       ipreflag=1)
    assert iret == 0 and hist[-1] < 1e-10
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-8


def test_cagmres_matches_gmres_of_same_restart():
    # same Krylov spaces and restart points as GMRES(s): histories agree to rounding
    rp, col, val, A = convdiff3d(6, peclet=1.0)
    b = ph.uniform(A.shape[0], 24)
    _, hc, nc, _ = orc.cagmres(rp, col, val, b, itrmax=41, rtol=1e-30, s=5, basis=1, qrflag=5, ipreflag=1)
    _, hg, ng, _ = orc.gmres(rp, col, val, b, itrmax=41, rtol=1e-30, m=5, ipreflag=1)
    assert nc == ng == 41
    np.testing.assert_allclose(hc, hg, rtol=1e-6)


@pytest.mark.parametrize("itrmax,s,niter,iret,written", [
    (1, 4, 1, 1, 5),   # the DO loop ends after one full cycle: itr = 5 >= itrmax
    (5, 4, 5, 0, 5),   # itr0 = i + itr reaches itrmax at i = 4
    (6, 4, 6, 0, 6),   # second cycle, i = 1
    (0, 4, 0, 1, 0),
])
def test_cagmres_exit_semantics(itrmax, s, niter, iret, written):
    rp, col, val, A = convdiff3d(4, peclet=1.0)
    n = A.shape[0]
    b = ph.uniform(n, 25)
    hist = np.full(itrmax + s + 1, -1.0)
    it, ret = ctypes.c_int(0), ctypes.c_int(0)
    x = np.zeros(n)
    rc = orc.load().oracle_cagmres(n, orc._p(rp), orc._p(col), orc._p(val), orc._p(b), orc._p(x), 0, 1, 1, 1,
                                   itrmax, 1e-30, s, 1, 0, 4, ctypes.byref(orc.TILED), orc._p(hist),
                                   ctypes.byref(it), ctypes.byref(ret))
    assert rc == 0
    assert (it.value, ret.value) == (niter, iret)
    assert int(np.sum(hist != -1.0)) == written


def test_cagmres_converged_at_restart_check():
    rp, col, val, A = convdiff3d(4)
    xs = ph.uniform(A.shape[0], 26)
    b = orc.spmv(rp, col, val, xs)
    x, hist, niter, iret = orc.cagmres(rp, col, val, b, x0=xs, itrmax=50, rtol=1e-8, s=4)
    assert (niter, iret) == (1, 0)
    assert np.array_equal(x, xs)


# --- t > 1: block CGS between blocks, Hessenberg from the block R factors ----
#
# The minimal-residual pins need a basis whose change-of-basis matrix is
# exact.  The reference's Newton recurrence (calc_NewtonBasisVec,
# pcagmres.f90:1480-1552) adds B1(i-1,i) V_{i-1} = -Im^2 V_{i-1} for a
# conjugate shift pair while its Hessenberg formulas use the same B1, which
# describes +Im^2; the oracle keeps the reference's sign, so Newton-basis
# pins use a symmetric operator (Poisson with point Jacobi: constant
# diagonal, real Ritz values), and the monomial basis covers the
# nonsymmetric problem.

def _spd_pj():
    rp, col, val = ph.poisson3d_spd(7, 6, 6)
    n = rp.shape[0] - 1
    A = orc_csr(rp, col, val)
    return rp, col, val, A


def orc_csr(rp, col, val):
    import scipy.sparse as sp
    n = rp.shape[0] - 1
    return sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))


BLOCK_CASES = [(0, (3, 2)), (0, (2, 4)), (0, (1, 5)), (1, (3, 2)), (1, (2, 4)), (1, (4, 3)), (1, (1, 5))]


@pytest.mark.parametrize("qr", [1, 2, 4, 5])
@pytest.mark.parametrize("basis,st", BLOCK_CASES)
def test_cagmres_blocks_first_cycle_is_minimal_residual(basis, st, qr):
    # one cycle of s*t steps over t blocks spans the Krylov space of
    # dimension s*t (whatever the per-block basis): every history entry is
    # the minimal residual of GMRES at that step
    s, t = st
    if basis == 0:
        rp, col, val, A = convdiff3d(6, peclet=2.0)
        pre = dict(ipreflag=2, ilu_method=0, nblocks=4)
        minv = _precon(rp, col, val, 2, ilu_method=0, nblocks=4)
    else:
        rp, col, val, A = _spd_pj()
        pre = dict(ipreflag=1)
        minv = _precon(rp, col, val, 1)
    b = ph.uniform(A.shape[0], 31)
    m = s * t
    x, hist, niter, iret = orc.cagmres(rp, col, val, b, itrmax=m + 1, rtol=1e-30, s=s, t=t, basis=basis,
                                       qrflag=qr, **pre)
    assert niter == m + 1 and iret == 0
    ref = min_residuals(A, minv, b, m) / np.linalg.norm(b)
    np.testing.assert_allclose(hist, ref, rtol=1e-6, atol=0)
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert abs(true - hist[-1]) <= 1e-6 * hist[-1]


@pytest.mark.parametrize("qr", [2, 4, 5])
@pytest.mark.parametrize("basis", [0, 1])
def test_cagmres_blocks_solution_at_every_exit_point(qr, basis):
    # stopping at step j of a cycle (itrmax = j + 1) exercises the solution
    # update from a block before the last (orthonormal Q) and from the last
    # block (its V columns times R^-1, for every count of them)
    if basis == 0:
        rp, col, val, A = convdiff3d(5, peclet=2.0)
    else:
        rp, col, val, A = _spd_pj()
    b = ph.uniform(A.shape[0], 32)
    s, t = 3, 3
    for j in range(1, s * t + 1):
        x, hist, niter, iret = orc.cagmres(rp, col, val, b, itrmax=j + 1, rtol=1e-30, s=s, t=t, basis=basis,
                                           qrflag=qr, ipreflag=1)
        assert niter == j + 1 and len(hist) == j + 1
        true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert abs(true - hist[-1]) <= 1e-7 * hist[-1], (qr, j)


@pytest.mark.parametrize("basis,qr", [(0, 4), (0, 5), (0, 1)])
def test_cagmres_blocks_match_gmres_of_same_restart(basis, qr):
    rp, col, val, A = convdiff3d(6, peclet=1.0)
    b = ph.uniform(A.shape[0], 33)
    s, t = 3, 2
    _, hc, nc, _ = orc.cagmres(rp, col, val, b, itrmax=37, rtol=1e-30, s=s, t=t, basis=basis, qrflag=qr,
                               ipreflag=1)
    _, hg, ng, _ = orc.gmres(rp, col, val, b, itrmax=37, rtol=1e-30, m=s * t, ipreflag=1)
    assert nc == ng == 37
    np.testing.assert_allclose(hc, hg, rtol=1e-6)


def test_cagmres_blocks_converge():
    rp, col, val, A = _spd_pj()
    b = ph.uniform(A.shape[0], 34)
    x, hist, niter, iret = orc.cagmres(rp, col, val, b, itrmax=3000, rtol=1e-10, s=4, t=5, basis=1, qrflag=5,
                                       ipreflag=1)
    assert iret == 0 and hist[-1] < 1e-10
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-8


def test_cagmres_unsupported_options():
    rp, col, val, A = convdiff3d(3)
    b = np.ones(A.shape[0])
    for kw in (dict(qrflag=6), dict(s=0)):
        with pytest.raises(RuntimeError):
            orc.cagmres(rp, col, val, b, itrmax=5, **kw)
