"""GPU tests of the Fortran module PARCEL_wrapper and the by-reference C
entry points: a Fortran program (parcel-legacy_amd/fortran/example_poisson.f90)
calls parcel_dcg / parcel_dbicgstab / parcel_dgmres exactly as the
reference's drivers do, and its printed histories and solutions must equal
the oracle's bit for bit."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "parcel-legacy_amd", "bin", "example_poisson")


def run_fortran(*args):
    print("example_poisson", *args, flush=True)
    out = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.split("\n")
    head = lines[0].split()
    niter, iret = int(head[1]), int(head[3])
    hist = np.array([float(v) for v in lines[1:1 + niter]])
    x1 = float(lines[1 + niter].split()[1])
    xn = float(lines[2 + niter].split()[1])
    return niter, iret, hist, x1, xn


@pytest.mark.parametrize("solver,g,ipre,ilu,nblk,itrmax,rtol,m,gs", [
    (1, 16, 2, 0, 4, 500, 1e-9, 4, 1),
    (1, 12, 1, 1, 1, 500, 1e-9, 4, 1),
    (2, 12, 2, 1, 6, 500, 1e-9, 4, 1),
    (3, 12, 2, 2, 3, 300, 1e-9, 8, 2),
    (3, 12, 0, 1, 1, 300, 1e-9, 10, 1),
])
def test_fortran_module_matches_oracle(solver, g, ipre, ilu, nblk, itrmax, rtol, m, gs):
    niter, iret, hist, x1, xn = run_fortran(solver, g, ipre, ilu, nblk, itrmax, rtol, m, gs)
    rp, col, val = ph.poisson3d_parcel(g, g, g)
    b, _ = ph.poisson3d_rhs_parcel(g, g, g)
    kw = dict(ipreflag=ipre, ilu_method=ilu, nblocks=nblk, itrmax=itrmax, rtol=rtol)
    if solver == 1:
        xo, ho, no, io = orc.pcg(rp, col, val, b, **kw)
    elif solver == 2:
        xo, ho, no, io = orc.bicgstab(rp, col, val, b, **kw)
    else:
        xo, ho, no, io = orc.gmres(rp, col, val, b, m=m, gsflag=gs, **kw)
    assert (niter, iret) == (no, io)
    assert np.array_equal(hist, ho)
    assert x1 == xo[0] and xn == xo[-1]


def test_by_reference_c_entry_point():
    lib = ph.load()
    g = 10
    rp, col, val = ph.poisson3d_parcel(g, g, g)
    b, _ = ph.poisson3d_rhs_parcel(g, g, g)
    n, nnz = g ** 3, int(col.shape[0])
    x = np.zeros(n)
    hist = np.zeros(400)
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    niter, iret = ctypes.c_int(400), ctypes.c_int(-7)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    lib.parcel_init_c_(I(0))
    lib.parcel_dgmres_c_(I(0), P(x), P(b), I(n), I(n), I(nnz), I(1), P(val), P(rp), P(col), I(2), I(0), I(0),
                         I(4), ctypes.byref(niter), ctypes.byref(ctypes.c_double(1e-9)), P(hist), I(0), I(4), I(1),
                         I(1), I(2000), I(6), I(1), ctypes.byref(iret))
    xo, ho, no, io = orc.gmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=400, rtol=1e-9, m=6)
    assert (niter.value, iret.value) == (no, io)
    assert np.array_equal(hist[:no], ho) and np.array_equal(x, xo)
    # a bad argument comes back as a negative iret instead of a stop
    niter = ctypes.c_int(10)
    lib.parcel_dgmres_c_(I(0), P(x), P(b), I(n), I(n), I(nnz), I(1), P(val), P(rp), P(col), I(2), I(0), I(0),
                         I(4), ctypes.byref(niter), ctypes.byref(ctypes.c_double(1e-9)), P(hist), I(0), I(4), I(1),
                         I(1), I(2000), I(0), I(1), ctypes.byref(iret))
    assert iret.value == -1
