"""GPU parity tests of the mixed-precision refinement (FP64 outer, FP32
PCG inner; no reference counterpart): outer history, solution, niter, iret
and the summed FP32 iteration count bit-exact against the oracle, which
runs the same FP32 arithmetic (float vectors and factors, fmaf, double dot
products in the tiled order)."""
import sys

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph
from problems import convdiff3d

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


PRECONS = {
    "none": dict(ipreflag=0),
    "pj": dict(ipreflag=1),
    "ilu0_8": dict(ipreflag=2, ILU_method=0, precon_thblock=8),
    "ilud_5": dict(ipreflag=2, ILU_method=1, precon_thblock=5),
    "mlud_3": dict(ipreflag=2, ILU_method=2, precon_thblock=3),
}


@pytest.mark.parametrize("pname", sorted(PRECONS))
@pytest.mark.parametrize("case", ["spd3d_24", "poisson2d_200"])
def test_mixed_bit_exact(case, pname):
    say("mixed", case, pname)
    if case == "spd3d_24":
        rp, col, val = ph.poisson3d_spd(24, 24, 24)
    else:
        rp, col, val = ph.poisson2d(200, 200, -1.0)
    n = rp.shape[0] - 1
    kw = PRECONS[pname]
    xt = ph.uniform(n, 7)
    b = orc.spmv(rp, col, val, xt)
    p = ph.default_param(solver=11, ipreflag=kw["ipreflag"], ILU_method=kw.get("ILU_method", 1),
                         precon_thblock=kw.get("precon_thblock", 1))
    plan = ph.Plan(rp, col, val, param=p)
    try:
        plan.set_inner(1e-4, 5000)
        x = np.zeros(n)
        hist, niter, iret = plan.solve(x, b, 20, 1e-10)
        st = plan.stats()
    finally:
        plan.close()
    nb = kw.get("precon_thblock", 1)
    if st["bj_lscan"]:
        # ILU(0) on blocks of whole grid lines of a symmetric matrix: the FP32
        # inner apply is the block IC(0) factor in line-scan form
        assert case == "poisson2d_200" and pname == "ilu0_8"
        nx = 200
        F64, F32 = orc.derive(*orc.ic0_grid(rp, col, val, nx, nblocks=nb), nx)
        M = orc.FactorPrecon(F64, F32, nx, nstrip=nb)
        xo, ho, no, io, inner = orc.pcg_mixed_factor(rp, col, val, b, M, itrmax=20, rtol=1e-10, inner_itrmax=5000,
                                                     inner_rtol=1e-4)
    else:
        xo, ho, no, io, inner = orc.pcg_mixed(rp, col, val, b, ipreflag=kw["ipreflag"],
                                              ilu_method=kw.get("ILU_method", 1), nblocks=nb, itrmax=20,
                                              rtol=1e-10, inner_itrmax=5000, inner_rtol=1e-4)
    assert (niter, iret) == (no, io)
    assert st["inner_niter"] == inner
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)
    assert st["final_rel_res"] < 1e-10


def test_mixed_one_shot_entry_and_large():
    say("mixed 2d 760")
    rp, col, val = ph.poisson2d(760, 760, -1.0)
    n = rp.shape[0] - 1
    b = ph.uniform(n, 8)
    x = np.zeros(n)
    hist, niter, iret = ph.dcg_mixed(x, b, rp, col, val, inner_rtol=1e-3, inner_itrmax=3000, ipreflag=2,
                                     ILU_method=0, precon_thblock=16, itrmax=30, rtolmax=1e-9)
    xo, ho, no, io, _ = orc.pcg_mixed(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=16, itrmax=30,
                                      rtol=1e-9, inner_itrmax=3000, inner_rtol=1e-3)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)
    with pytest.raises(ph.ParcelError):
        ph.dcg_mixed(np.zeros(n), b, rp, col, val, inner_rtol=0.0)
