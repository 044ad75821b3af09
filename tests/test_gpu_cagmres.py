"""GPU parity tests of CA-GMRES(s, t=1) (src/pcagmres.f90:1557-2517) through
the C ABI: residual history, solution, niter and iret bit-exact against the
oracle (the host-side dense algebra follows the same operation order and the
device reductions the oracle's tiled order)."""
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
    "parcel3d_20": lambda: ph.poisson3d_parcel(20, 20, 20),
    "poisson2d_760": lambda: ph.poisson2d(760, 760, -1.0),  # n = 577600 > 256*2048
}

PRECONS = {
    "none": dict(ipreflag=0),
    "pj": dict(ipreflag=1),
    "ilu0_4": dict(ipreflag=2, ILU_method=0, precon_thblock=4),
    "ilud_8": dict(ipreflag=2, ILU_method=1, precon_thblock=8),
}


def _both(rp, col, val, b, s, basis, qr, kw, itrmax, rtol, x0=None, t=1):
    n = rp.shape[0] - 1
    x = np.zeros(n) if x0 is None else x0.copy()
    hist, niter, iret = ph.dcagmres(x, b, rp, col, val, CAGMRES_sstep=s, CAGMRES_tstep=t, CAGMRES_basis=basis,
                                    CAGMRES_QRflag=qr, itrmax=itrmax, rtolmax=rtol, ipreflag=kw["ipreflag"],
                                    ILU_method=kw.get("ILU_method", 1), precon_thblock=kw.get("precon_thblock", 1))
    xo, ho, no, io = orc.cagmres(rp, col, val, b, x0=x0, ipreflag=kw["ipreflag"],
                                 ilu_method=kw.get("ILU_method", 1), nblocks=kw.get("precon_thblock", 1),
                                 itrmax=itrmax, rtol=rtol, s=s, t=t, basis=basis, qrflag=qr)
    return (x, hist, niter, iret), (xo, ho, no, io)


@pytest.mark.parametrize("qr", [1, 2, 4, 5])
@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("pname", ["pj", "ilu0_4"])
@pytest.mark.parametrize("case", ["convdiff_12", "parcel3d_20"])
def test_cagmres_bit_exact(case, pname, basis, qr):
    say("cagmres", case, pname, basis, qr)
    rp, col, val = PROBLEMS[case]()
    b = ph.uniform(rp.shape[0] - 1, 41)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 6, basis, qr, PRECONS[pname], 300, 1e-9)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho)
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("s,qr", [(1, 4), (2, 5), (8, 5), (8, 1), (15, 4), (15, 2)])
def test_cagmres_step_counts(s, qr):
    say("cagmres s", s, qr)
    rp, col, val = PROBLEMS["convdiff_12"]()
    b = ph.uniform(rp.shape[0] - 1, 42)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, s, 1, qr, PRECONS["ilud_8"], 200, 1e-10)
    assert (ni, ir) == (no, io)
    # s = 15 with CholQR breaks down (the monomial Gram matrix is numerically
    # singular): NaN in both, at the same place
    assert np.array_equal(h, ho, equal_nan=True) and np.array_equal(x, xo, equal_nan=True)


def test_cagmres_large_capped_tiling():
    say("cagmres poisson2d_760")
    rp, col, val = PROBLEMS["poisson2d_760"]()
    b = ph.uniform(rp.shape[0] - 1, 43)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 8, 1, 5, PRECONS["ilu0_4"], 33, 1e-12)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)


@pytest.mark.parametrize("itrmax", [0, 1, 5, 6, 13])
def test_cagmres_exit_semantics(itrmax):
    say("cagmres itrmax", itrmax)
    rp, col, val = PROBLEMS["convdiff_12"]()
    b = ph.uniform(rp.shape[0] - 1, 44)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 4, 1, 4, PRECONS["none"], itrmax, 1e-30)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)


def test_cagmres_converged_at_restart_and_errors():
    rp, col, val = PROBLEMS["convdiff_12"]()
    n = rp.shape[0] - 1
    xs = ph.uniform(n, 45)
    b = orc.spmv(rp, col, val, xs)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 4, 1, 4, PRECONS["pj"], 50, 1e-8, x0=xs)
    assert (ni, ir) == (no, io) == (1, 0)
    assert np.array_equal(x, xs)
    for kw, code in ((dict(CAGMRES_sstep=0), -1), (dict(CAGMRES_sstep=16), -1), (dict(CAGMRES_tstep=0), -1),
                     (dict(CAGMRES_sstep=8, CAGMRES_tstep=17), -1), (dict(CAGMRES_QRflag=6), -1)):
        with pytest.raises(ph.ParcelError) as e:
            ph.dcagmres(np.zeros(n), b, rp, col, val, itrmax=10, **kw)
        assert e.value.code == code


# --- t > 1: block CGS between blocks -------------------------------------------

@pytest.mark.parametrize("qr", [1, 2, 4, 5])
@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("s,t", [(3, 2), (4, 5), (1, 6), (8, 3)])
def test_cagmres_blocks_bit_exact(s, t, basis, qr):
    say("cagmres blocks", s, t, basis, qr)
    rp, col, val = PROBLEMS["convdiff_12"]()
    b = ph.uniform(rp.shape[0] - 1, 46)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, s, basis, qr, PRECONS["ilu0_4"], 250, 1e-9, t=t)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho, equal_nan=True) and np.array_equal(x, xo, equal_nan=True)


@pytest.mark.parametrize("itrmax", [2, 4, 7, 9, 10, 23])
@pytest.mark.parametrize("qr", [2, 5])
def test_cagmres_blocks_exit_points(itrmax, qr):
    # exits inside a block before the last (orthonormal Q update) and inside
    # the last block (V columns times R^-1), and at cycle ends
    say("cagmres blocks itrmax", itrmax, qr)
    rp, col, val = PROBLEMS["parcel3d_20"]()
    b = ph.uniform(rp.shape[0] - 1, 47)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 3, 1, qr, PRECONS["pj"], itrmax, 1e-30, t=3)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)


def test_cagmres_blocks_large():
    say("cagmres blocks poisson2d_760")
    rp, col, val = PROBLEMS["poisson2d_760"]()
    b = ph.uniform(rp.shape[0] - 1, 48)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 4, 1, 5, PRECONS["ilu0_4"], 50, 1e-12, t=4)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)
