"""GPU parity tests of the TSQR R factor (TSQR_Ronly_,
src/QRDecomposition.f90:887-1002) through the C ABI, and of CA-GMRES with
CAGMRES_QRflag = 3: bit-exact against the oracle (same panels per
wavefront, same butterfly sums, same host Householder on the stacked
factors)."""
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


@pytest.mark.parametrize("n,nc", [(65, 1), (100, 3), (1000, 5), (4097, 9), (70000, 16), (1 << 20, 5),
                                  (3_000_001, 4)])
def test_tsqr_r_bit_exact(n, nc):
    say("tsqr", n, nc)
    rng = np.random.default_rng(n + 7 * nc)
    V = rng.standard_normal((n, nc))
    R = ph.tsqr_r(V)
    Ro = orc.tsqr_r(V)
    assert np.array_equal(R, Ro)
    assert np.all(np.diag(R) > 0)


def test_tsqr_r_zero_column():
    rng = np.random.default_rng(2)
    V = rng.standard_normal((5000, 4))
    V[:, 1] = 0.0
    assert np.array_equal(ph.tsqr_r(V), orc.tsqr_r(V))


def _both(rp, col, val, b, s, t, basis, itrmax, rtol, ipreflag=1):
    n = rp.shape[0] - 1
    x = np.zeros(n)
    h, ni, ir = ph.dcagmres(x, b, rp, col, val, CAGMRES_sstep=s, CAGMRES_tstep=t, CAGMRES_basis=basis,
                            CAGMRES_QRflag=3, itrmax=itrmax, rtolmax=rtol, ipreflag=ipreflag, precon_thblock=1)
    xo, ho, no, io = orc.cagmres(rp, col, val, b, ipreflag=ipreflag, nblocks=1, itrmax=itrmax, rtol=rtol, s=s,
                                 t=t, basis=basis, qrflag=3)
    return (x, h, ni, ir), (xo, ho, no, io)


@pytest.mark.parametrize("s,t,basis", [(4, 1, 1), (8, 1, 0), (3, 3, 1), (4, 5, 1), (1, 4, 0)])
def test_cagmres_tsqr_bit_exact(s, t, basis):
    say("cagmres tsqr", s, t, basis)
    rp, col, val = convdiff3d(12, peclet=2.0)[:3]
    b = ph.uniform(rp.shape[0] - 1, 51)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, s, t, basis, 300, 1e-9)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)


def test_cagmres_tsqr_short_vectors_use_cholqr2():
    rp, col, val = convdiff3d(4, peclet=1.0)[:3]  # n = 64
    b = ph.uniform(64, 52)
    (x, h, ni, ir), (xo, ho, no, io) = _both(rp, col, val, b, 4, 2, 1, 40, 1e-12)
    assert (ni, ir) == (no, io)
    assert np.array_equal(h, ho) and np.array_equal(x, xo)
