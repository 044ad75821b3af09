"""GPU parity tests of restarted GMRES(m) (src/pgmres.f90:62-525) through the
C ABI: residual history, solution, niter and iret bit-exact against the
oracle, which uses the GPU's reduction order."""
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
    "convdiff_12": lambda: convdiff3d(12, peclet=2.0)[:3],
    "parcel3d_24": lambda: ph.poisson3d_parcel(24, 24, 24),
    "poisson2d_800": lambda: ph.poisson2d(800, 800, -1.0),  # n = 640000 > 256*2048: capped tiling
}

PRECONS = {
    "none": dict(ipreflag=0),
    "pj": dict(ipreflag=1),
    "ilu0_4": dict(ipreflag=2, ILU_method=0, precon_thblock=4),
    "ilud_8": dict(ipreflag=2, ILU_method=1, precon_thblock=8),
    "mlud_3": dict(ipreflag=2, ILU_method=2, precon_thblock=3),
}


def _both(rp, col, val, b, m, gs, kw, itrmax, rtol, x0=None):
    n = rp.shape[0] - 1
    x = np.zeros(n) if x0 is None else x0.copy()
    hist, niter, iret = ph.dgmres(x, b, rp, col, val, GMRES_m=m, GMRES_GSflag=gs, itrmax=itrmax, rtolmax=rtol,
                                  ipreflag=kw["ipreflag"], ILU_method=kw.get("ILU_method", 1),
                                  precon_thblock=kw.get("precon_thblock", 1))
    xo, ho, no, io = orc.gmres(rp, col, val, b, x0=x0, ipreflag=kw["ipreflag"],
                               ilu_method=kw.get("ILU_method", 1), nblocks=kw.get("precon_thblock", 1),
                               itrmax=itrmax, rtol=rtol, m=m, gsflag=gs)
    return (x, hist, niter, iret), (xo, ho, no, io)


@pytest.mark.parametrize("gs", [1, 2])
@pytest.mark.parametrize("pname", sorted(PRECONS))
@pytest.mark.parametrize("case", ["convdiff_12", "parcel3d_24"])
def test_gmres_bit_exact(case, pname, gs):
    say("gmres", case, pname, gs)
    rp, col, val = PROBLEMS[case]()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 31)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 12, gs, PRECONS[pname], 400, 1e-9)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho)
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("m,gs", [(1, 1), (1, 2), (9, 2), (17, 2), (30, 1)])
def test_gmres_restart_lengths(m, gs):
    # CGS with more than 8 columns spans several dot-product launches
    say("gmres m", m, gs)
    rp, col, val = PROBLEMS["convdiff_12"]()
    b = ph.uniform(rp.shape[0] - 1, 32)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, m, gs, PRECONS["pj"], 120, 1e-12)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)


def test_gmres_large_capped_tiling():
    say("gmres poisson2d_800")
    rp, col, val = PROBLEMS["poisson2d_800"]()
    b = ph.uniform(rp.shape[0] - 1, 33)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 6, 1, PRECONS["ilu0_4"], 20, 1e-12)
    assert (ni, ir) == (no, io) == (20, 0)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)


@pytest.mark.parametrize("itrmax", [0, 1, 4, 6, 9])
def test_gmres_exit_semantics(itrmax):
    say("gmres itrmax", itrmax)
    rp, col, val = PROBLEMS["convdiff_12"]()
    b = ph.uniform(rp.shape[0] - 1, 34)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 4, 1, PRECONS["none"], itrmax, 1e-30)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)


def test_gmres_converged_at_restart_check():
    rp, col, val = PROBLEMS["convdiff_12"]()
    n = rp.shape[0] - 1
    xs = ph.uniform(n, 35)
    b = orc.spmv(rp, col, val, xs)  # r = b - A xs = 0 exactly
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 5, 1, PRECONS["pj"], 50, 1e-8, x0=xs)
    assert (ni, ir) == (no, io) == (2, 0)
    assert np.array_equal(x, xs) and np.array_equal(x, xo)


def test_gmres_plan_reuse_and_errors():
    rp, col, val = PROBLEMS["convdiff_12"]()
    n = rp.shape[0] - 1
    b = ph.uniform(n, 36)
    p = ph.default_param(solver=3, ipreflag=2, ILU_method=0, precon_thblock=4, GMRES_m=8, GMRES_GSflag=2)
    plan = ph.Plan(rp, col, val, param=p)
    try:
        x1 = np.zeros(n)
        r1 = plan.solve(x1, b, 200, 1e-10)
        x2 = np.zeros(n)
        r2 = plan.solve(x2, b, 200, 1e-10)
        assert r1[1:] == r2[1:] and np.array_equal(r1[0], r2[0]) and np.array_equal(x1, x2)
        xo, ho, no, io = orc.gmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=200, rtol=1e-10,
                                   m=8, gsflag=2)
        assert np.array_equal(r1[0], ho) and np.array_equal(x1, xo) and r1[1:] == (no, io)
    finally:
        plan.close()
    for bad_m in (0, 257):
        x = np.zeros(n)
        with pytest.raises(ph.ParcelError) as e:
            ph.dgmres(x, b, rp, col, val, GMRES_m=bad_m, itrmax=10)
        assert e.value.code == -1
