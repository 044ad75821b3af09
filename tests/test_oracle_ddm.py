"""CPU tests of the stencil (DDM) path's problem and box numbering: the
diagonal-format generator against the CRS generator, and the box-major CRS
against the CRS matrix permuted (the GPU test compares the solver with the
oracle on this CRS bit for bit, and with the reference's golden history)."""
import numpy as np
import scipy.sparse as sp

import parcel_hip as ph
from problems import csr_to_scipy, ddm_box_crs, ddm_box_split


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
