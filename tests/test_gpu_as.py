"""GPU parity of the CRS additive Schwarz preconditioner (ipreflag 3;
AdditiveSchwarz_crs at one rank, src/matrix.f90:3809-4375,4749-4858) through
the C ABI: the preconditioner apply and whole PCG solves bit for bit against
the oracle, for every overlap mode (iflagAS 1 BASIC, 2 RESTRICT, 3
INTERPOLATE, 4 NONE), both D-ILU forms, independ_nvec > 1, and the cases the
reference turns into block Jacobi."""
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
    "poisson2d_40": lambda: ph.poisson2d(40, 40, -1.0),
    "parcel3d_16": lambda: ph.poisson3d_parcel(16, 16, 16),
    "convdiff_12": lambda: convdiff3d(12, peclet=1.0)[:3],
}


@pytest.mark.parametrize("flag", [1, 2, 3, 4])
@pytest.mark.parametrize("ilu", [1, 2])
@pytest.mark.parametrize("case,nb,addL,nvec", [("poisson2d_40", 4, 5, 1), ("parcel3d_16", 8, 100, 1),
                                                ("convdiff_12", 6, 17, 1), ("parcel3d_16", 5, 40, 4)])
def test_as_apply_bit_exact(case, nb, addL, nvec, ilu, flag):
    say("AS apply", case, nb, addL, nvec, ilu, flag)
    rp, col, val = PROBLEMS[case]()
    n = rp.shape[0] - 1
    r = ph.uniform(n, 5)
    with ph.Plan(rp, col, val, ipreflag=3, ILU_method=ilu, precon_thblock=nb, addL=addL, iflagAS=flag,
                 independ_nvec=nvec) as P:
        s = np.zeros(n)
        P.precon(r, s)
    M = orc.AsPrecon(rp, col, val, ilu_method=ilu, nblocks=nb, addL=addL, iflagAS=flag, independ_nvec=nvec)
    assert np.array_equal(s, M.apply(r))


@pytest.mark.parametrize("flag", [1, 2, 3, 4])
@pytest.mark.parametrize("case,nb,addL", [("poisson2d_40", 4, 5), ("parcel3d_16", 8, 100)])
def test_as_pcg_bit_exact(case, nb, addL, flag):
    say("AS pcg", case, nb, addL, flag)
    rp, col, val = PROBLEMS[case]()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 9)
    x = np.zeros(n)
    hist, niter, iret = ph.dcg(x, b, rp, col, val, ipreflag=3, ILU_method=1, precon_thblock=nb, addL=addL,
                               iflagAS=flag, itrmax=1000, rtolmax=1e-9)
    M = orc.AsPrecon(rp, col, val, ilu_method=1, nblocks=nb, addL=addL, iflagAS=flag)
    xo, ho, no, io = orc.pcg_factor(rp, col, val, b, M, itrmax=1000, rtol=1e-9)
    assert (niter, iret) == (no, io)
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)


def test_as_large_3d_pcg():
    # 64^3 Poisson, 64 blocks of 4096 rows (one grid plane each), half a
    # plane of overlap (check_AS_region turns a wider overlap than half a
    # block into block Jacobi)
    say("AS pcg 64^3")
    g = 64
    rp, col, val = ph.poisson3d_spd(g, g, g)
    n = g ** 3
    b = orc.spmv(rp, col, val, ph.uniform(n, 10))
    x = np.zeros(n)
    hist, niter, iret = ph.dcg(x, b, rp, col, val, ipreflag=3, ILU_method=1, precon_thblock=64,
                               addL=g * g // 2, iflagAS=1, itrmax=1000, rtolmax=1e-10)
    M = orc.AsPrecon(rp, col, val, ilu_method=1, nblocks=64, addL=g * g // 2, iflagAS=1)
    xo, ho, no, io = orc.pcg_factor(rp, col, val, b, M, itrmax=1000, rtol=1e-10)
    assert (niter, iret) == (no, io) and iret == 0
    assert np.array_equal(hist, ho) and np.array_equal(x, xo)
    # the overlap pays: fewer iterations than block Jacobi on the same blocks
    _, hb, nbj, _ = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=1, nblocks=64, itrmax=1000, rtol=1e-10)
    say("iterations AS", niter, "BJ", nbj)
    assert niter < nbj


@pytest.mark.parametrize("ilu", [0, 1, 2])
def test_as_degenerate_cases_are_block_jacobi(ilu):
    rp, col, val = ph.poisson2d(30, 30, -1.0)
    n = rp.shape[0] - 1
    r = ph.uniform(n, 6)
    bj = orc.Precon(rp, col, val, 2, ilu_method=1 if ilu == 1 else 2, nblocks=6).apply(r)
    for addL in (0, 76):  # no overlap; more than half of a 150-row block
        with ph.Plan(rp, col, val, ipreflag=3, ILU_method=ilu, precon_thblock=6, addL=addL, iflagAS=2) as P:
            s = np.zeros(n)
            P.precon(r, s)
        assert np.array_equal(s, bj), addL
