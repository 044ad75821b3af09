"""Long rows in the SELL-64 layout, on the GPU through the C ABI, bit for bit
against the oracle.

A slice of 64 rows is as wide as its rows after the few longest; those keep
the rest of their entries in an overflow area and the same lane sums them
after the slice part, so every row is still summed in CRS order
(MatVec_crs_FP64_CPU, src/matrix.f90:9051-9110).  The matrices are the 3D
Poisson stencil with a few hub rows coupled to hundreds of random columns
(symmetric and diagonally dominant, or with the couplings in the hub rows
only), which without the overflow would pad their slices to the hub length.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import parcel_hip as ph

pytestmark = pytest.mark.gpu

HUBS = ((7, 300), (1000, 1000), (1001, 40), (2049, 150), (4000, 64))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


def hub_matrix(g=16, sym=True, seed=0):
    rp, col, val = ph.poisson3d_spd(g, g, g)
    n = rp.shape[0] - 1
    A = sp.csr_matrix((val, col - 1, rp - 1), shape=(n, n)).tolil()
    rng = np.random.default_rng(seed)
    for h, m in HUBS:
        js = rng.choice(np.setdiff1d(np.arange(n), [h]), size=m, replace=False)
        w = -0.01 * (1.0 + rng.random(m))
        for j, a in zip(js, w):
            A[h, j] += a
            if sym:
                A[j, h] += a
                A[j, j] += -a
        A[h, h] += -w.sum() + 1.0
    A = A.tocsr()
    A.sort_indices()
    return (A.indptr.astype(np.int32) + 1, A.indices.astype(np.int32) + 1, A.data.astype(np.float64))


@pytest.mark.parametrize("sym", [True, False])
def test_long_rows_spmv_and_layout(sym):
    say("long rows spmv, sym", sym)
    rp, col, val = hub_matrix(sym=sym)
    n = rp.shape[0] - 1
    lens = np.diff(rp)
    assert lens.max() > 1000
    x = ph.uniform(n, 5)
    with ph.Plan(rp, col, val, ipreflag=0) as P:
        y = np.zeros(n)
        P.spmv(x, y)
        st = P.stats()
    assert np.array_equal(y, orc.spmv(rp, col, val, x))
    # the hub slices are as wide as their other rows: the stored entries stay
    # near nnz, where padding every slice to its longest row would store 4.4 nnz
    assert st["sell_overflow"] > 0
    padded = sum(64 * int(lens[s:s + 64].max()) for s in range(0, n, 64))
    assert st["sell_entries"] < 1.5 * lens.sum() and padded > 4 * lens.sum()


@pytest.mark.parametrize("pre", [dict(ipreflag=1), dict(ipreflag=2, ILU_method=0, precon_thblock=4),
                                 dict(ipreflag=2, ILU_method=1, precon_thblock=3)])
def test_long_rows_pcg(pre):
    say("long rows pcg", pre)
    rp, col, val = hub_matrix()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 9)
    x = np.zeros(n)
    hist, niter, iret = ph.dcg(x, b, rp, col, val, itrmax=2000, rtolmax=1e-10, **pre)
    xo, ho, no, io = orc.pcg(rp, col, val, b, ipreflag=pre["ipreflag"], ilu_method=pre.get("ILU_method", 1),
                             nblocks=pre.get("precon_thblock", 1), itrmax=2000, rtol=1e-10)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)


def test_long_rows_nonsymmetric_solvers():
    rp, col, val = hub_matrix(sym=False, seed=1)
    n = rp.shape[0] - 1
    b = ph.uniform(n, 11)
    say("long rows bicgstab")
    x = np.zeros(n)
    hist, niter, iret = ph.dbicgstab(x, b, rp, col, val, itrmax=2000, rtolmax=1e-10, ipreflag=1)
    xo, ho, no, io = orc.bicgstab(rp, col, val, b, ipreflag=1, itrmax=2000, rtol=1e-10)
    assert (niter, iret) == (no, io) and np.array_equal(hist, ho) and np.array_equal(x, xo)
    say("long rows gmres")
    x = np.zeros(n)
    hist, niter, iret = ph.dgmres(x, b, rp, col, val, GMRES_m=20, GMRES_GSflag=1, itrmax=2000, rtolmax=1e-10,
                                  ipreflag=2, ILU_method=0, precon_thblock=4)
    xo, ho, no, io = orc.gmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=2000, rtol=1e-10,
                               m=20, gsflag=1)
    assert (niter, iret) == (no, io) and np.array_equal(hist, ho) and np.array_equal(x, xo)


def test_long_rows_dd_and_mixed():
    rp, col, val = hub_matrix(seed=2)
    n = rp.shape[0] - 1
    b = ph.uniform(n, 13)
    say("long rows ddcg")
    vlo = val * 2.0 ** -58
    blo = b * 2.0 ** -57
    xh, xl = np.zeros(n), np.zeros(n)
    hist, niter, iret = ph.ddcg(xh, xl, b, blo, rp, col, val, vlo, ipreflag=1, itrmax=3000, rtolmax=1e-20,
                                PRECISION_A=2, PRECISION_b=2, PRECISION_x=2, PRECISION_PRECONDITION=2)
    oh, ol, ho, no, io = orc.ddcg(rp, col, val, vlo, b, blo, ipreflag=1, prec=2, itrmax=3000, rtol=1e-20)
    assert (niter, iret) == (no, io)
    assert np.array_equal(hist, ho) and np.array_equal(xh, oh) and np.array_equal(xl, ol)
    say("long rows mixed")
    x = np.zeros(n)
    hist, niter, iret = ph.dcg_mixed(x, b, rp, col, val, inner_rtol=1e-4, inner_itrmax=3000, ipreflag=2,
                                     ILU_method=0, precon_thblock=4, itrmax=30, rtolmax=1e-10)
    xo, ho, no, io, _ = orc.pcg_mixed(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=30,
                                      rtol=1e-10, inner_itrmax=3000, inner_rtol=1e-4)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)
