"""CPU tests of the oracle: pinned against the reference's golden outputs,
closed forms, textbook routines and brute force (no GPU needed)."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

import oracle as orc
import parcel_hip as ph
from problems import convdiff3d

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "templates.json")))


def csr_to_scipy(rp, col, val):
    n = rp.shape[0] - 1
    return sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))


def random_csr(n, nnz_per_row, seed, diag=10.0):
    """1-based CSR with sorted columns and a dominant diagonal."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        k = int(rng.integers(0, nnz_per_row + 1))
        cs = set(int(c) for c in rng.integers(0, n, size=k))
        cs.add(i)
        for c in sorted(cs):
            rows.append(i)
            cols.append(c)
            vals.append(diag if c == i else rng.uniform(-1, 1))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return (A.indptr.astype(np.int32) + 1, A.indices.astype(np.int32) + 1, A.data.astype(np.float64))


# --- reductions --------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 63, 256, 1000, 70001, 600000])
def test_dot_orders_vs_exact(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    exact = math.fsum(float(a) * float(b) for a, b in zip(x.tolist(), y.tolist())) if n < 100000 else \
        float(np.dot(x.astype(np.longdouble), y.astype(np.longdouble)))
    scale = float(np.abs(x) @ np.abs(y)) + 1e-300
    for red in (orc.TILED, orc.SEQ):
        d = orc.dot(x, y, red)
        assert abs(d - exact) <= 4 * n * 2.0**-53 * scale


def test_tiled_dot_small_exact_integers():
    # integer data: every order gives the exact sum
    x = np.arange(1, 3001, dtype=np.float64)
    assert orc.dot(x, x, orc.TILED) == float(sum(i * i for i in range(1, 3001)))


def test_tiled_dot_is_the_documented_tree():
    # brute-force restatement of the tiled order for a size spanning several tiles
    n, nt, nbmax = 5000, 256, 2048
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    ntiles = -(-n // nt)
    nb = min(ntiles, nbmax)
    parts = []
    for b in range(nb):
        acc = [0.0] * nt
        for k in range(b, ntiles, nb):
            for t in range(nt):
                i = k * nt + t
                if i < n:
                    acc[t] = math.fma(x[i], y[i], acc[t]) if hasattr(math, "fma") else \
                        float(Fraction(x[i]) * Fraction(y[i]) + Fraction(acc[t]))
        tot = None
        for w in range(nt // 64):
            v = acc[64 * w: 64 * w + 64]
            o = 32
            while o >= 1:
                v = [v[l] + v[l ^ o] for l in range(64)]
                o //= 2
            tot = v[0] if tot is None else tot + v[0]
        parts.append(tot)
    assert orc.dot(x, y, orc.TILED) == orc.reduce_partials(np.array(parts))


# --- divide_region (src/common.f90:125-182) -----------------------------------

@pytest.mark.parametrize("n,np_", [(10, 3), (51200 * 80, 80), (7, 7), (1000, 256), (65536, 256), (5, 1)])
def test_divide_region_partition(n, np_):
    spans = [orc.divide_region(r, np_, n) for r in range(np_)]
    assert spans[0][0] == 1 and spans[-1][1] == n
    for (a, b), (c, d) in zip(spans, spans[1:]):
        assert c == b + 1
    sizes = [b - a + 1 for a, b in spans]
    assert max(sizes) - min(sizes) <= 1
    # the first mod(n, np) blocks are the larger ones
    over = n % np_
    assert all(s == n // np_ + 1 for s in sizes[:over])


# --- SpMV ------------------------------------------------------------------------

def test_spmv_matches_scipy_random():
    rp, col, val = random_csr(3000, 12, 1)
    x = np.random.default_rng(2).standard_normal(3000)
    y = orc.spmv(rp, col, val, x)
    ref = csr_to_scipy(rp, col, val) @ x
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-13)


def test_spmv_poisson_closed_form():
    # A * ones has row sums: 0 in the interior of the parcel 3D operator
    rp, col, val = ph.poisson3d_parcel(12, 9, 7)
    y = orc.spmv(rp, col, val, np.ones(12 * 9 * 7))
    A = csr_to_scipy(rp, col, val)
    assert np.array_equal(y, np.asarray(A.sum(axis=1)).ravel())
    assert (A != A.T).nnz == 0
    nbr = np.asarray((A != 0).sum(axis=1)).ravel()
    assert y[np.where(nbr == 7)[0]].max() == 0.0


# --- preconditioners -------------------------------------------------------------

def test_point_jacobi_closed_form():
    rp, col, val = random_csr(500, 6, 3)
    A = csr_to_scipy(rp, col, val)
    r = np.random.default_rng(4).standard_normal(500)
    s = orc.Precon(rp, col, val, 1).apply(r)
    assert np.array_equal(s, r * (1.0 / A.diagonal()))


def tridiag(n, a, b, c):
    rows, cols, vals = [], [], []
    for i in range(n):
        if i > 0:
            rows.append(i); cols.append(i - 1); vals.append(b)
        rows.append(i); cols.append(i); vals.append(a)
        if i < n - 1:
            rows.append(i); cols.append(i + 1); vals.append(c)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return A.indptr.astype(np.int32) + 1, A.indices.astype(np.int32) + 1, A.data.astype(np.float64), A


def test_ilu0_tridiagonal_is_exact_lu():
    # ILU(0) of a tridiagonal matrix has no dropped fill: it is the exact LU,
    # so one block solves A s = r to rounding.
    rp, col, val, A = tridiag(400, 4.0, -1.3, -0.7)
    r = np.random.default_rng(5).standard_normal(400)
    s = orc.Precon(rp, col, val, 2, ilu_method=0, nblocks=1).apply(r)
    P, L, U = scipy.linalg.lu(A.toarray())
    assert np.allclose(P, np.eye(400))
    ref = scipy.linalg.solve_triangular(U, scipy.linalg.solve_triangular(L, r, lower=True, unit_diagonal=True))
    assert np.allclose(s, ref, rtol=1e-13, atol=1e-14)


def test_ilu0_blocks_are_independent():
    # with k blocks the result equals per-block exact solves of the diagonal blocks
    rp, col, val, A = tridiag(300, 4.0, -1.0, -1.0)
    r = np.random.default_rng(6).standard_normal(300)
    s = orc.Precon(rp, col, val, 2, ilu_method=0, nblocks=7).apply(r)
    Ad = A.toarray()
    for b in range(7):
        lo, hi = orc.divide_region(b, 7, 300)
        blk = Ad[lo - 1:hi, lo - 1:hi]
        assert np.allclose(s[lo - 1:hi], np.linalg.solve(blk, r[lo - 1:hi]), rtol=1e-12, atol=1e-14)


def test_dilu_1d_continued_fraction():
    # D-ILU of tridiag(b, a, c): D_1 = a, D_k = a - b*c / D_{k-1} (exact fractions)
    n = 40
    rp, col, val, A = tridiag(n, 4.0, -1.0, -0.5)
    D = [Fraction(4)]
    for _ in range(n - 1):
        D.append(Fraction(4) - Fraction(-1) * Fraction(-1, 2) / D[-1])
    # K = (D+L) D^-1 (D+U): check K s = r for the oracle's s
    r = np.random.default_rng(8).standard_normal(n)
    s = orc.Precon(rp, col, val, 2, ilu_method=1, nblocks=1).apply(r)
    Dm = np.diag([float(d) for d in D])
    L = np.tril(A.toarray(), -1)
    U = np.triu(A.toarray(), 1)
    K = (Dm + L) @ np.linalg.inv(Dm) @ (Dm + U)
    assert np.allclose(K @ s, r, rtol=1e-12, atol=1e-12)


def test_dilu_mlud_row_sum_property():
    # MLUD keeps row sums: K * ones == A * ones
    rp, col, val = ph.poisson3d_parcel(6, 5, 4)
    A = csr_to_scipy(rp, col, val).toarray()
    n = A.shape[0]
    P = orc.Precon(rp, col, val, 2, ilu_method=2, nblocks=1)
    ones = np.ones(n)
    s = P.apply(A @ ones)
    assert np.allclose(s, ones, rtol=1e-10)


# --- the reference's golden residual histories -------------------------------------

def _golden(name, block=0, column=1):
    return np.array(GOLD[name]["histories"][block])[:, column]


@pytest.mark.slow
def test_pcg_dilu_80_blocks_matches_reference_cpu_template():
    """template_cpu/STDOUT_run: 160^3 Poisson, CG + block-Jacobi D-ILU, 4 ranks x
    20 threads = 80 contiguous blocks, rtol 1e-9, 291 iterations."""
    rp, col, val = ph.poisson3d_parcel(160, 160, 160)
    b, xe = ph.poisson3d_rhs_parcel(160, 160, 160)
    x, hist, niter, iret = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=1, nblocks=80, itrmax=2000,
                                   rtol=1e-9)
    gold = _golden("template_cpu/STDOUT_run")
    assert niter == len(gold) == 291 and iret == 0
    assert np.max(np.abs(hist - gold) / gold) <= 1e-12
    err = np.abs(x - xe).max()
    assert abs(err - GOLD["template_cpu/STDOUT_run"]["extras"]["max_abs_err"][0]) <= 1e-6 * err


@pytest.mark.slow
def test_pcg_ilu0_4_blocks_matches_reference_gpu_template():
    """template_gpu/STDOUT_run: same problem, cuSPARSE ILU(0) with one block per
    rank (4 ranks), 227 iterations; both printed columns."""
    rp, col, val = ph.poisson3d_parcel(160, 160, 160)
    b, xe = ph.poisson3d_rhs_parcel(160, 160, 160)
    x, hist, niter, iret = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=2000,
                                   rtol=1e-9)
    for c in (1, 2):
        gold = _golden("template_gpu/STDOUT_run", column=c)
        assert niter == len(gold) == 227 and iret == 0
        assert np.max(np.abs(hist - gold) / gold) <= 1e-12


def test_planted_mistake_is_caught():
    """A plausible slip (dropping the off-block test so D-ILU sums across block
    boundaries == one block) must move the history far outside tolerance on a
    problem where the reference golden pins it; checked on a small analogue:
    80 blocks vs 1 block differ at the 1e-3 level by iteration 3."""
    rp, col, val = ph.poisson3d_parcel(20, 20, 20)
    b, _ = ph.poisson3d_rhs_parcel(20, 20, 20)
    _, h80, _, _ = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=1, nblocks=80, itrmax=5, rtol=1e-30)
    _, h1, _, _ = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=1, nblocks=1, itrmax=5, rtol=1e-30)
    assert np.max(np.abs(h80[2:] - h1[2:]) / h1[2:]) > 1e-3


def test_pcg_exit_semantics():
    # itrmax reached: iret = 1, niter = itrmax (pcg.f90:850-855)
    rp, col, val = ph.poisson2d(16, 16, -1.0)
    b = ph.uniform(256, 0)
    _, hist, niter, iret = orc.pcg(rp, col, val, b, ipreflag=1, itrmax=5, rtol=1e-12)
    assert niter == 5 and iret == 1 and hist[0] == 1.0
    # converged: history holds the pre-update ratios, the last one >= rtol
    x, hist, niter, iret = orc.pcg(rp, col, val, b, ipreflag=1, itrmax=500, rtol=1e-10)
    assert iret == 0 and hist[-1] >= 1e-10
    A = csr_to_scipy(rp, col, val)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9


# --- BiCGStab (src/pbicg.f90:280-698) --------------------------------------------

def test_bicgstab_diagonal_one_iteration():
    # A = diag(d), no preconditioner: r0 = b, q = A b, alpha = <b,b>/<b,Ab>,
    # e = b - alpha A b, v = A e, c3 = <e,v>/<v,v>; with a point-Jacobi
    # preconditioner M^-1 A = I and the first iteration is exact
    n = 50
    d = np.linspace(1.0, 3.0, n)
    A = sp.diags(d).tocsr()
    rp, col, val = A.indptr.astype(np.int32) + 1, A.indices.astype(np.int32) + 1, A.data.copy()
    b = np.random.default_rng(3).standard_normal(n)
    x, hist, niter, iret = orc.bicgstab(rp, col, val, b, ipreflag=1, itrmax=10, rtol=1e-14)
    assert niter == 1 and iret == 0
    assert np.allclose(x, b / d, rtol=1e-15, atol=0)


def test_bicgstab_matches_scipy_nonsymmetric():
    import scipy.sparse.linalg as spla
    rp, col, val, A = convdiff3d(10)
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    x, hist, niter, iret = orc.bicgstab(rp, col, val, b, ipreflag=0, itrmax=500, rtol=1e-10)
    assert iret == 0
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9
    its = []
    xs, info = spla.bicgstab(A, b, rtol=1e-10, atol=0, maxiter=500, callback=lambda xk: its.append(1))
    assert info == 0
    assert abs(len(its) - niter) <= 2
    assert np.allclose(x, xs, rtol=1e-6, atol=1e-9 * np.abs(xs).max())


def test_bicgstab_block_jacobi_converges_faster():
    rp, col, val, A = convdiff3d(12)
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    _, _, n0, _ = orc.bicgstab(rp, col, val, b, ipreflag=0, itrmax=500, rtol=1e-10)
    x, _, n2, iret = orc.bicgstab(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=500, rtol=1e-10)
    assert iret == 0 and n2 < n0
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9
