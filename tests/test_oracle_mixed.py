"""CPU tests of the oracle's mixed-precision iterative refinement (FP64
outer correction around an FP32 PCG; no reference counterpart).  Pinned by
what the arithmetic fixes: the returned x solves the system to the FP64
tolerance, the recorded history is the true FP64 residual before each
correction, each correction contracts the residual by about the inner
tolerance, and the inner solve is genuinely FP32 (it stalls near FP32
round-off when asked for more, while the outer loop still converges)."""
import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph
from problems import csr_to_scipy

PRECONS = [dict(ipreflag=1), dict(ipreflag=2, ilu_method=0, nblocks=8), dict(ipreflag=2, ilu_method=1, nblocks=8)]


@pytest.mark.parametrize("kw", PRECONS, ids=["pj", "ilu0", "dilu"])
def test_mixed_refinement_converges(kw):
    g = 20
    rp, col, val = ph.poisson3d_spd(g, g, g)
    A = csr_to_scipy(rp, col, val)
    xt = ph.uniform(g ** 3, 3)
    b = A @ xt
    x, hist, niter, iret, inner = orc.pcg_mixed(rp, col, val, b, rtol=1e-10, inner_rtol=1e-4, **kw)
    assert iret == 0 and 2 <= niter <= 4 and inner > 0
    assert hist[0] == 1.0                                   # x0 = 0
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert true < 1e-10
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-8
    # each correction contracts the FP64 residual by about the inner tolerance
    ratios = np.append(hist[1:], true) / hist
    assert np.all(ratios < 5e-4)


def test_mixed_history_is_true_residual_before_each_correction():
    rp, col, val = ph.poisson3d_spd(12, 12, 12)
    A = csr_to_scipy(rp, col, val)
    b = ph.uniform(12 ** 3, 5)
    for k in (1, 2):
        xk, hk, nk, ik, _ = orc.pcg_mixed(rp, col, val, b, itrmax=k, rtol=1e-30, ipreflag=1)
        assert (nk, ik) == (k, 1)
        x2, h2, _, _, _ = orc.pcg_mixed(rp, col, val, b, itrmax=k + 1, rtol=1e-30, ipreflag=1)
        assert np.array_equal(h2[:k], hk)
        assert abs(h2[k] - np.linalg.norm(b - A @ xk) / np.linalg.norm(b)) <= 1e-12 * h2[k]


def test_inner_solve_is_single_precision():
    # asking the FP32 inner PCG for 1e-12 cannot be met: its recurrence
    # residual drifts from the true one at FP32 round-off, so the first
    # correction lands near 1e-7, not 1e-12; refinement still reaches 1e-12
    rp, col, val = ph.poisson3d_spd(16, 16, 16)
    A = csr_to_scipy(rp, col, val)
    b = ph.uniform(16 ** 3, 6)
    x, hist, niter, iret, inner = orc.pcg_mixed(rp, col, val, b, rtol=1e-12, inner_rtol=1e-12,
                                                inner_itrmax=400, ipreflag=1)
    assert iret == 0 and niter >= 2
    assert hist[1] > 1e-9                                   # FP32 floor after one correction
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-12
