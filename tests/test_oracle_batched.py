"""CPU tests of the random-coefficient generator and of the oracle's factor
preconditioners (IC(0) on the five-point pattern, the line-scan solves).

The generator is pinned by an independent numpy restatement (numpy's FFT),
its circulant eigenvalues by an FFT of the covariance, and its statistics.
The oracle's IC(0) and scan solves are pinned by the mathematics (exact
Cholesky of a tridiagonal, L L^T = A on the pattern, scipy triangular
solves), so the bit-exact GPU tests compare against something checked."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle as orc
import parcel_hip as ph


def test_normal_stream_matches_numpy_restatement():
    x = ph.normal(4097, 11)
    xo = orc.normal(4097, 11)
    assert np.max(np.abs(x - xo)) < 1e-15
    # Box-Muller of a 53-bit uniform stream: moments of N(0,1)
    y = ph.normal(200000, 3)
    assert abs(y.mean()) < 0.01 and abs(y.std() - 1.0) < 0.01


def test_grf_eigenvalues_closed_form_equal_fft_of_covariance():
    for nx, ny, ell in ((512, 512, 16.0), (256, 128, 8.0)):
        i, j = np.arange(nx), np.arange(ny)
        di, dj = np.minimum(i, nx - i), np.minimum(j, ny - j)
        c = 4.0 * np.exp(-(dj[:, None] ** 2 + di[None, :] ** 2) / (2 * ell * ell))
        lam = np.fft.fft2(c).real
        lc = orc.grf_eigenvalues(nx, ny, ell, 2.0)
        assert np.max(np.abs(lam - lc)) < 1e-12 * lam.max()
        assert lc.min() >= 0.0  # exp underflows to 0 only below 1e-300


@pytest.mark.parametrize("nx,ny", [(64, 64), (512, 512), (128, 256)])
def test_grf_matches_numpy_restatement(nx, ny):
    g = ph.grf2d(nx, ny, 16.0, 2.0, 1000)
    go = orc.grf2d(nx, ny, 16.0, 2.0, 1000)
    assert np.max(np.abs(g - go)) < 1e-12


def test_grf_statistics():
    # variance sigma^2 and correlation exp(-d^2/(2 ell^2)) over 16 fields
    gs = np.stack([ph.grf2d(512, 512, 16.0, 2.0, 1000 + s).reshape(512, 512) for s in range(16)])
    var = gs.var()
    assert abs(var - 4.0) < 0.4
    for d in (8, 16, 32):
        cov = np.mean(gs * np.roll(gs, d, axis=2))
        assert abs(cov / 4.0 - np.exp(-d * d / 512.0)) < 0.06


def test_diffusion2d_matches_numpy_and_is_spd():
    nx, ny = 32, 16
    kap = np.exp(ph.grf2d(32, 16, 4.0, 2.0, 7))
    rp, col, val = ph.diffusion2d(nx, ny, kap)
    rpo, colo, valo = orc.diffusion2d(nx, ny, kap)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)
    assert np.array_equal(val, valo)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(nx * ny,) * 2).toarray()
    assert np.allclose(A, A.T, rtol=0, atol=0)
    assert np.linalg.eigvalsh(A).min() > 0
    # constant kappa: the 4 / -1 Laplacian plus kappa per Dirichlet face
    rp, col, val = ph.diffusion2d(nx, ny, np.full(nx * ny, 3.0))
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(nx * ny,) * 2).toarray()
    assert np.allclose(np.diag(A), 12.0) and np.allclose(A[1, 0], -3.0)


def test_batch_generator_matches_single():
    rp, col, vals, g = ph.grf_diffusion2d(64, 32, 1000, 3, ell=8.0, fields=True)
    for s in range(3):
        gs = ph.grf2d(64, 32, 8.0, 2.0, 1000 + s)
        assert np.array_equal(g[s], gs)
        r1, c1, v1 = ph.diffusion2d(64, 32, np.exp(gs))  # numpy's exp: within an ulp of libm's
        assert np.array_equal(r1, rp) and np.array_equal(c1, col)
        np.testing.assert_allclose(v1, vals[s], rtol=2e-15, atol=0)


def five_point(nx, ny, seed=5):
    kap = np.exp(ph.grf2d(128, 128, 4.0, 1.5, seed)[: nx * ny]) if nx * ny <= 128 * 128 else \
        np.exp(ph.grf2d(1024, 1024, 4.0, 1.5, seed)[: nx * ny])
    return ph.diffusion2d(nx, ny, kap)


def lower_factor(Ls, Lw, Ld, nx):
    n = Ld.shape[0]
    i = np.arange(n)
    rows = np.concatenate([i, i[nx:], i[i % nx > 0]])
    cols = np.concatenate([i, i[nx:] - nx, i[i % nx > 0] - 1])
    vals = np.concatenate([Ld, Ls[nx:], Lw[i % nx > 0]])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_ic0_is_exact_cholesky_on_a_tridiagonal():
    nx = 50
    rp, col, val = five_point(nx, 1)
    Ls, Lw, Ld = orc.ic0_5pt(rp, val, nx)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(nx, nx)).toarray()
    C = sla.cholesky(A, lower=True)
    L = lower_factor(Ls, Lw, Ld, nx).toarray()
    assert np.allclose(L, C, rtol=1e-13, atol=1e-13)


def test_ic0_reproduces_a_on_its_pattern():
    nx, ny = 40, 30
    rp, col, val = five_point(nx, ny)
    n = nx * ny
    Ls, Lw, Ld = orc.ic0_5pt(rp, val, nx)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))
    L = lower_factor(Ls, Lw, Ld, nx)
    D = (L @ L.T - A).multiply(A != 0)
    assert abs(D).max() < 1e-12 * abs(A).max()


# line layouts: one wavefront (nx <= 128), two (128 < nx <= 256), four
# (nx > 256: E = 2 or 4 cells per lane, 300 leaves a whole wavefront of padding)
@pytest.mark.parametrize("nx,ny", [(64, 8), (100, 37), (200, 6), (256, 4), (300, 5), (512, 4), (1000, 3), (1, 50)])
def test_lscan_solves_llt(nx, ny):
    rp, col, val = five_point(nx, ny)
    n = nx * ny
    Ls, Lw, Ld = orc.ic0_5pt(rp, val, nx)
    F64, F32 = orc.derive(Ls, Lw, Ld, nx)
    r = np.random.default_rng(1).standard_normal(n)
    L = lower_factor(Ls, Lw, Ld, nx)
    ref = spl.spsolve_triangular(L.T.tocsr(), spl.spsolve_triangular(L, r, lower=True), lower=False)
    z = orc.lscan_apply(F64, r, nx)
    assert np.max(np.abs(z - ref)) < 1e-12 * np.max(np.abs(ref))
    z32 = orc.lscan_apply(F32, r.astype(np.float32), nx)
    assert np.max(np.abs(z32 - ref)) < 1e-4 * np.max(np.abs(ref))


def test_lscan_plants_a_transposed_solve_and_fails():
    # the check above catches a mistake: solving L L^T with the forward and
    # backward factors swapped (L^T L) does not reproduce the reference
    nx, ny = 100, 20
    rp, col, val = five_point(nx, ny)
    Ls, Lw, Ld = orc.ic0_5pt(rp, val, nx)
    F64, _ = orc.derive(Ls, Lw, Ld, nx)
    r = np.random.default_rng(2).standard_normal(nx * ny)
    L = lower_factor(Ls, Lw, Ld, nx)
    wrong = spl.spsolve_triangular(L, spl.spsolve_triangular(L.T.tocsr(), r, lower=False), lower=True)
    z = orc.lscan_apply(F64, r, nx)
    assert np.max(np.abs(z - wrong)) > 1e-3 * np.max(np.abs(z))


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5e-3, 3.0e38, np.inf], np.float32)
    h = orc.f32_to_bf16(x)
    back = orc.bf16_to_f32(h)
    assert back[0] == 1.0
    assert back[1] == 1.0                       # tie -> even
    assert back[2] == np.float32(1.0078125)     # above the tie -> up
    assert np.all(np.abs(back[:5] - x[:5]) <= np.abs(x[:5]) * 2.0 ** -8)
    assert np.isinf(back[5])
    for v in np.random.default_rng(0).standard_normal(1000).astype(np.float32):
        assert int(orc.f32_to_bf16(np.array([v]))[0]) == orc.f32_to_bf16_c(v)


def test_factor_pcg_oracle_converges_faster_than_jacobi():
    nx, ny = 64, 64
    rp, col, val = five_point(nx, ny)
    n = nx * ny
    xt = np.random.default_rng(3).standard_normal(n)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))
    b = A @ xt
    M = orc.FactorPrecon(*orc.derive(*orc.ic0_5pt(rp, val, nx), nx), nx)
    x, h, it, iret = orc.pcg_factor(rp, col, val, b, M, itrmax=2000, rtol=1e-10)
    _, _, itj, _ = orc.pcg(rp, col, val, b, ipreflag=1, itrmax=2000, rtol=1e-10)
    assert iret == 0 and it < 0.6 * itj
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10


@pytest.mark.parametrize("nx,ny,nstrip", [(100, 37, 4), (64, 16, 16), (512, 9, 2)])
def test_lscan_strips_solve_block_diagonal(nx, ny, nstrip):
    # strips drop the factor's couplings between line blocks: the apply equals
    # solving diag(L_k L_k^T) block by block
    rp, col, val = five_point(nx, ny)
    n = nx * ny
    Ls, Lw, Ld = orc.ic0_5pt(rp, val, nx)
    F64, _ = orc.derive(Ls, Lw, Ld, nx)
    r = np.random.default_rng(4).standard_normal(n)
    z = orc.lscan_apply(F64, r, nx, nstrip=nstrip)
    ref = np.zeros(n)
    for k in range(nstrip):
        y0, y1 = ny * k // nstrip, ny * (k + 1) // nstrip
        sl = slice(y0 * nx, y1 * nx)
        Lk = lower_factor(Ls[sl].copy(), Lw[sl].copy(), Ld[sl].copy(), nx)
        Lk = sp.csr_matrix(Lk)
        ref[sl] = spl.spsolve_triangular(Lk.T.tocsr(), spl.spsolve_triangular(Lk, r[sl], lower=True), lower=False)
    assert np.max(np.abs(z - ref)) < 1e-12 * np.max(np.abs(ref))


def test_block_ic0_is_ic0_of_each_diagonal_block():
    # block-Jacobi IC(0) over 3 strips of grid lines: each strip's factor is
    # the IC(0) of the strip's diagonal block alone, and L L^T equals the
    # block-diagonal part of A on the pattern
    import parcel_hip as ph
    nx, ny, nb = 12, 9, 3
    rp, col, val = ph.diffusion2d(nx, ny, np.exp(ph.grf2d(16, 16, 3.0, 1.0, 8)[: nx * ny]))
    n = nx * ny
    Ls, Lw, Ld = orc.ic0_5pt(rp, val, nx, nblocks=nb)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n)).toarray()
    L = np.diag(Ld) + np.diag(Lw[1:], -1) + np.diag(Ls[nx:], -nx)
    for k in range(nb):
        a, b = ny * k // nb * nx, ny * (k + 1) // nb * nx
        Ab = A[a:b, a:b]
        Ab = Ab.copy()
        sub = sp.csr_matrix(Ab)
        sub.sort_indices()
        S = orc.ic0_5pt(sub.indptr.astype(np.int32) + 1, sub.data.copy(), nx)
        assert np.array_equal(Ls[a:b][nx:], S[0][nx:]) and np.all(Ls[a:a + nx] == 0)
        assert np.array_equal(Lw[a:b], S[1]) and np.array_equal(Ld[a:b], S[2])
    LLt = L @ L.T
    Abd = np.zeros_like(A)
    for k in range(nb):
        a, b = ny * k // nb * nx, ny * (k + 1) // nb * nx
        Abd[a:b, a:b] = A[a:b, a:b]
    pattern = A != 0
    assert np.allclose(LLt[pattern & (Abd != 0)], Abd[pattern & (Abd != 0)], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nb", [1, 3, 7])
def test_ic0_grid_equals_ic0_5pt_on_five_point_matrices(nb):
    # the column-search restatement reproduces the positional one bit for bit
    import parcel_hip as ph
    nx, ny = 14, 21
    rp, col, val = ph.diffusion2d(nx, ny, np.exp(ph.grf2d(32, 32, 4.0, 1.5, 3)[: nx * ny]))
    a = orc.ic0_grid(rp, col, val, nx, nblocks=nb)
    b = orc.ic0_5pt(rp, val, nx, nblocks=nb)
    for u, v in zip(a, b):
        assert u.tobytes() == v.tobytes()


def test_ic0_grid_on_3d_plane_blocks_is_ic0_of_each_plane():
    # 3D seven-point Poisson, one block per grid plane: the plane couplings
    # (offset nx ny) are outside every block, and each plane's factor is the
    # IC(0) of the plane's own five-point matrix; L L^T equals A's in-plane
    # entries on the pattern
    import parcel_hip as ph
    nx, ny, nz = 9, 7, 5
    rp, col, val = ph.poisson3d_spd(nx, ny, nz)
    n = nx * ny * nz
    Ls, Lw, Ld = orc.ic0_grid(rp, col, val, nx, nblocks=nz)
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n))
    for k in range(nz):
        a, b = k * nx * ny, (k + 1) * nx * ny
        sub = sp.csr_matrix(A[a:b, a:b])
        sub.sort_indices()
        S = orc.ic0_5pt(sub.indptr.astype(np.int32) + 1, sub.data.copy(), nx)
        assert np.array_equal(Ls[a:b], S[0]) and np.array_equal(Lw[a:b], S[1]) and np.array_equal(Ld[a:b], S[2])
        m = b - a
        L = np.diag(Ld[a:b]) + np.diag(Lw[a + 1:b], -1) + np.diag(Ls[a + nx:b], -nx)
        Ad = sub.toarray()
        pat = Ad != 0
        assert np.allclose((L @ L.T)[pat], Ad[pat], rtol=1e-12, atol=1e-12)
    # a planted mistake: strips one line off the planes mix two planes
    Ls2, _, _ = orc.ic0_grid(rp, col, val, nx, nblocks=nz + 1)
    assert not np.array_equal(Ls2, Ls)
