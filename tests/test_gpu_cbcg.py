"""GPU parity tests of Chebyshev-basis CG (KSP_pCBCG, src/pcbcg.f90:1182-2025)
through the C ABI: residual history, solution, niter and iret bit-exact
against the oracle (device dots in the tiled order, host pseudo-inverse and
small GEMMs in the oracle's operation order), with the power-method bound
(src/power_method.f90:188-437) and with the fixed bound of an unknown
CBCG_Eigenflag."""
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


PROBLEMS = {
    "spd3d_20": lambda: ph.poisson3d_spd(20, 20, 20),
    "parcel3d_16": lambda: ph.poisson3d_parcel(16, 16, 16),
    "convdiff_10": lambda: convdiff3d(10, peclet=0.3)[:3],
    "poisson2d_760": lambda: ph.poisson2d(760, 760, -1.0),  # n = 577600 > 256*2048
}

PRECONS = {
    "none": dict(ipreflag=0),
    "pj": dict(ipreflag=1),
    "ilu0_4": dict(ipreflag=2, ILU_method=0, precon_thblock=4),
    "ilud_8": dict(ipreflag=2, ILU_method=1, precon_thblock=8),
}


def _both(rp, col, val, b, k, eig, pkw, itrmax, rtol, x0=None, pm=30):
    n = rp.shape[0] - 1
    x = np.zeros(n) if x0 is None else x0.copy()
    hist, niter, iret = ph.dcbcg(x, b, rp, col, val, CBCG_kstep=k, CBCG_Eigenflag=eig, power_method_itrmax=pm,
                                 itrmax=itrmax, rtolmax=rtol, ipreflag=pkw["ipreflag"],
                                 ILU_method=pkw.get("ILU_method", 1), precon_thblock=pkw.get("precon_thblock", 1))
    xo, ho, no, io, _ = orc.cbcg(rp, col, val, b, x0=x0, ipreflag=pkw["ipreflag"],
                                 ilu_method=pkw.get("ILU_method", 1), nblocks=pkw.get("precon_thblock", 1),
                                 itrmax=itrmax, rtol=rtol, k=k, eigflag=eig, pm_itrmax=pm)
    return (x, hist, niter, iret), (xo, ho, no, io)


def _same(g, o):
    x, h, n, r = g
    xo, ho, no, io = o
    assert (n, r) == (no, io)
    assert np.array_equal(h, ho)
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("k", [1, 2, 4, 8])
@pytest.mark.parametrize("pname", ["none", "pj", "ilu0_4", "ilud_8"])
@pytest.mark.parametrize("case", ["spd3d_20", "parcel3d_16", "convdiff_10"])
def test_cbcg_bit_exact(case, pname, k):
    say("cbcg", case, pname, k)
    rp, col, val = PROBLEMS[case]()
    b = ph.uniform(rp.shape[0] - 1, 43)
    g, o = _both(rp, col, val, b, k, 1, PRECONS[pname], itrmax=600, rtol=1e-8)
    _same(g, o)
    assert g[1].shape[0] >= 2


def test_cbcg_large_and_fixed_bound():
    rp, col, val = PROBLEMS["poisson2d_760"]()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 5)
    say("cbcg large", n)
    g, o = _both(rp, col, val, b, 4, 1, PRECONS["ilu0_4"], itrmax=80, rtol=1e-10, pm=10)
    _same(g, o)
    # flag 2 (CA-Arnoldi, the power method in the GPU build), the default
    g, o = _both(rp, col, val, b, 4, 2, PRECONS["ilud_8"], itrmax=80, rtol=1e-10, pm=10)
    _same(g, o)
    # an unknown eigenvalue flag uses Emax = 1, Emin = 0 (a valid Krylov basis)
    g, o = _both(rp, col, val, b, 3, 0, PRECONS["pj"], itrmax=60, rtol=1e-10)
    _same(g, o)


def test_cbcg_exit_semantics_and_start():
    rp, col, val = ph.poisson3d_spd(10, 10, 10)
    n = rp.shape[0] - 1
    b = ph.uniform(n, 7)
    # loop runs out: niter = itrmax/k + 1 with a zero last entry, iret 0
    g, o = _both(rp, col, val, b, 4, 1, PRECONS["pj"], itrmax=8, rtol=1e-30)
    _same(g, o)
    assert g[2] == 3 and g[3] == 0 and g[1][2] == 0.0
    # k = 1 at the limit: iret 1
    g, o = _both(rp, col, val, b, 1, 1, PRECONS["pj"], itrmax=5, rtol=1e-30)
    _same(g, o)
    assert g[2] == 5 and g[3] == 1
    # nonzero initial guess (also the power method's start vector)
    x0 = ph.uniform(n, 9)
    g, o = _both(rp, col, val, b, 4, 1, PRECONS["ilu0_4"], itrmax=400, rtol=1e-9, x0=x0)
    _same(g, o)


def test_cbcg_rejects_bad_k():
    rp, col, val = ph.poisson3d_spd(6, 6, 6)
    n = rp.shape[0] - 1
    with pytest.raises(ph.ParcelError):
        ph.dcbcg(np.zeros(n), np.ones(n), rp, col, val, CBCG_kstep=9, CBCG_Eigenflag=1)
