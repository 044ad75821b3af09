"""CPU tests of the stencil (DDM) path's problem and box numbering: the
diagonal-format generator against the CRS generator, and the box-major CRS
against the CRS matrix permuted (the GPU test compares the solver with the
oracle on this CRS bit for bit, and with the reference's golden history)."""
import numpy as np
import scipy.sparse as sp

import parcel_hip as ph
import oracle as orc
from problems import csr_to_scipy, ddm_box_crs, ddm_box_split, ddm_ras


def test_box_split_matches_reference_rule():
    # n/4 boxes of ceil(n / (n/4)) cells
    assert ddm_box_split(160) == list(range(0, 161, 4))
    assert ddm_box_split(40) == list(range(0, 41, 4))
    assert ddm_box_split(10) == [0, 5, 10]
    assert ddm_box_split(3) == [0, 3]
    assert ddm_box_split(13) == [0, 5, 10, 13]


def test_ddm_generator_is_the_crs_problem_padded():
    nx, ny, nz = 6, 5, 4
    d = ph.poisson3d_ddm(nx, ny, nz)
    b, xe = ph.poisson3d_rhs_parcel(nx, ny, nz)
    mx, my = nx + 2, ny + 2
    idx = np.array([(x + 1) + mx * (y + 1) + mx * my * (z + 1)
                    for z in range(nz) for y in range(ny) for x in range(nx)])
    assert np.array_equal(d["b"][idx], b) and np.array_equal(d["x_exact"][idx], xe)
    mask = np.ones(d["m"], bool)
    mask[idx] = False
    assert np.all(d["b"][mask] == 0) and np.all(d["val_dia"].reshape(7, -1)[:, mask] == 0)


def test_box_crs_is_the_permuted_matrix():
    nx, ny, nz = 9, 8, 7
    d = ph.poisson3d_ddm(nx, ny, nz)
    rp, col, val, perm, starts = ddm_box_crs(d)
    A = csr_to_scipy(*ph.poisson3d_parcel(nx, ny, nz))
    mx, my = nx + 2, ny + 2
    lin = lambda j: ((j % mx) - 1) + nx * ((j // mx) % my - 1) + nx * ny * (j // (mx * my) - 1)
    p = np.array([lin(j) for j in perm])
    B = csr_to_scipy(rp, col, val)
    assert abs(B - A[p][:, p]).max() == 0
    assert starts[0] == 0 and starts[-1] == nx * ny * nz and np.all(np.diff(starts) > 0)


def _ras(d, addL):
    rp, col, val, perm, starts = ddm_box_crs(d)
    src, dst, rpe, cole, vale, st = ddm_ras(d, addL)
    return (rp, col, val, perm, starts), orc.RasPrecon(perm.size, src, dst, rpe, cole, vale, st), (src, dst)


def test_ras_without_overlap_is_the_box_dilu():
    d = ph.poisson3d_ddm(13, 10, 9)
    (rp, col, val, perm, starts), R, (src, dst) = _ras(d, 0)
    assert np.array_equal(src, np.arange(perm.size)) and np.array_equal(dst, src)
    B = orc.Precon(rp, col, val, 2, ilu_method=1, starts=starts)
    r = np.random.default_rng(1).standard_normal(perm.size)
    assert np.array_equal(R.apply(r), B.apply(r))


def test_ras_on_a_line_is_restricted_exact_local_solves():
    # one grid line: each extended box matrix is tridiagonal, whose D-ILU is
    # its exact LU, so RAS = owned part of the exact solve on each extended box
    for n, addL in ((16, 1), (16, 2), (21, 1), (40, 3)):
        d = ph.poisson3d_ddm(n, 1, 1)
        (rp, col, val, perm, starts), R, (src, dst) = _ras(d, addL)
        A = csr_to_scipy(rp, col, val).toarray()
        r = np.random.default_rng(n).standard_normal(n)
        ex = ddm_box_split(n)
        lin = np.argsort(perm)  # box-major rows are the x order on a line
        s_ref = np.zeros(n)
        for k in range(len(ex) - 1):
            lo = ex[k] - addL if k > 0 else ex[k]
            hi = ex[k + 1] + addL if k < len(ex) - 2 else ex[k + 1]
            ext = lin[lo:hi]
            z = np.linalg.solve(A[np.ix_(ext, ext)], r[ext])
            own = lin[ex[k]:ex[k + 1]]
            s_ref[own] = z[ex[k] - lo: ex[k] - lo + len(own)]
        assert np.allclose(R.apply(r), s_ref, rtol=1e-13, atol=1e-13 * np.abs(s_ref).max()), (n, addL)


def test_ras_overlap_lowers_iterations():
    d = ph.poisson3d_ddm(16, 16, 16)
    rp, col, val, perm, starts = ddm_box_crs(d)
    b = d["b"][perm].copy()
    its = []
    for addL in (0, 1, 2):
        _, R, _ = _ras(d, addL)
        _, h, nit, ret = orc.pcg_factor(rp, col, val, b, R, itrmax=2000, rtol=1e-9)
        assert ret == 0
        its.append(nit)
    assert its[0] > its[1] > its[2], its

