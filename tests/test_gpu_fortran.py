"""GPU tests of the Fortran module PARCEL_wrapper and the by-reference C
entry points: a Fortran program (parcel-legacy_amd/fortran/example_poisson.f90)
calls every module solver exactly as the reference's drivers do, and a C
program in the form of the reference's GPU-build driver
(parcel-legacy_amd/fortran/example_devicemem.c) passes device arrays as
devicemem handles through libparcel_devicemem.so; the printed histories and
solutions must equal the oracle's bit for bit."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph
from problems import ddm_box_crs, ddm_ras

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "parcel-legacy_amd", "bin", "example_poisson")
EXE_DEV = os.path.join(ROOT, "parcel-legacy_amd", "bin", "example_devicemem")


def run_fortran(*args, exe=EXE):
    print(os.path.basename(exe), *args, flush=True)
    out = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.split("\n")
    head = lines[0].split()
    niter, iret = int(head[1]), int(head[3])
    hist = np.array([float(v) for v in lines[1:1 + niter]])
    x1 = float(lines[1 + niter].split()[1])
    xn = float(lines[2 + niter].split()[1])
    return niter, iret, hist, x1, xn


CASES = [
    (1, 16, 2, 0, 4, 500, 1e-9, (4, 1)),
    (1, 12, 1, 1, 1, 500, 1e-9, (4, 1)),
    (2, 12, 2, 1, 6, 500, 1e-9, (4, 1)),
    (3, 12, 2, 2, 3, 300, 1e-9, (8, 2)),
    (3, 12, 0, 1, 1, 300, 1e-9, (10, 1)),
    (4, 12, 2, 0, 4, 400, 1e-9, (4, 2, 1, 3)),
    (4, 12, 1, 1, 1, 400, 1e-9, (5, 1, 0, 5)),
    (5, 12, 2, 1, 3, 400, 1e-9, (4, 1)),
    (6, 10, 2, 0, 4, 400, 1e-20, ()),
    (7, 12, 2, 1, 1, 400, 1e-9, (0,)),
    (7, 12, 3, 1, 1, 400, 1e-9, (1,)),
]


def oracle_case(solver, g, ipre, ilu, nblk, itrmax, rtol, p):
    """(x1, xn, hist, niter, iret) of the oracle for a driver case."""
    if solver == 7:
        d = ph.poisson3d_ddm(g, g, g)
        rp, col, val, perm, starts = ddm_box_crs(d)
        if ipre == 3:
            src, dst, rpe, cole, vale, st = ddm_ras(d, p[0])
            M = orc.RasPrecon(perm.size, src, dst, rpe, cole, vale, st)
        else:
            M = orc.Precon(rp, col, val, ipre, ilu_method=1, starts=starts if ipre == 2 else None)
        xo, ho, no, io = orc.pcg_factor(rp, col, val, d["b"][perm].copy(), M, itrmax=itrmax, rtol=rtol)
        return xo[0], xo[-1], ho, no, io
    rp, col, val = ph.poisson3d_parcel(g, g, g)
    b, _ = ph.poisson3d_rhs_parcel(g, g, g)
    kw = dict(ipreflag=ipre, ilu_method=ilu, nblocks=nblk, itrmax=itrmax, rtol=rtol)
    if solver in (1, 8):
        xo, ho, no, io = orc.pcg(rp, col, val, b, **kw)
    elif solver == 2:
        xo, ho, no, io = orc.bicgstab(rp, col, val, b, **kw)
    elif solver in (3, 9):
        xo, ho, no, io = orc.gmres(rp, col, val, b, m=p[0], gsflag=p[1], **kw)
    elif solver == 4:
        xo, ho, no, io = orc.cagmres(rp, col, val, b, s=p[0], t=p[1], basis=p[2], qrflag=p[3], **kw)
    elif solver == 5:
        xo, ho, no, io, _ = orc.cbcg(rp, col, val, b, k=p[0], eigflag=p[1], pm_itrmax=100, **kw)
    else:
        z = np.zeros_like(b)
        xo, _, ho, no, io = orc.ddcg(rp, col, val, np.zeros_like(val), b, z, prec=2, **kw)
    return xo[0], xo[-1], ho, no, io


@pytest.mark.parametrize("case", CASES)
def test_fortran_module_matches_oracle(case):
    solver, g, ipre, ilu, nblk, itrmax, rtol, p = case
    niter, iret, hist, x1, xn = run_fortran(solver, g, ipre, ilu, nblk, itrmax, rtol, *p)
    o1, on, ho, no, io = oracle_case(*case)
    assert (niter, iret) == (no, io)
    assert np.array_equal(hist, ho)
    assert x1 == o1 and xn == on


# the DIA module procedures (diagonals from parcel_gen_crs_to_dia): DIA = CRS bit for bit
DIA_CASES = [
    (8, 12, 2, 0, 4, 500, 1e-9, (4, 1)),
    (8, 12, 1, 1, 1, 500, 1e-9, (4, 1)),
    (9, 12, 2, 1, 3, 300, 1e-9, (8, 2)),
]


@pytest.mark.parametrize("case", DIA_CASES)
def test_fortran_module_dia_matches_oracle(case):
    solver, g, ipre, ilu, nblk, itrmax, rtol, p = case
    niter, iret, hist, x1, xn = run_fortran(solver, g, ipre, ilu, nblk, itrmax, rtol, *p)
    o1, on, ho, no, io = oracle_case(*case)
    assert (niter, iret) == (no, io)
    assert np.array_equal(hist, ho)
    assert x1 == o1 and xn == on


@pytest.mark.parametrize("case", CASES)
def test_devicemem_c_driver_matches_oracle(case):
    # the reference GPU build's C calling convention: devicemem handles
    solver, g, ipre, ilu, nblk, itrmax, rtol, p = case
    niter, iret, hist, x1, xn = run_fortran(solver, g, ipre, ilu, nblk, itrmax, rtol, *p, exe=EXE_DEV)
    o1, on, ho, no, io = oracle_case(*case)
    assert (niter, iret) == (no, io)
    assert np.array_equal(hist, ho)
    assert x1 == o1 and xn == on


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


def test_by_reference_dcagmres_and_ddcg_entry_points():
    # the host-array by-reference forms of parcel_dcagmres_c and parcel_ddcg_c
    lib = ph.load()
    g = 10
    rp, col, val = ph.poisson3d_parcel(g, g, g)
    b, _ = ph.poisson3d_rhs_parcel(g, g, g)
    n, nnz = g ** 3, int(col.shape[0])
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    D = lambda v: ctypes.byref(ctypes.c_double(v))
    lib.parcel_init_c_(I(0))
    x, hist = np.zeros(n), np.zeros(500)
    niter, iret = ctypes.c_int(400), ctypes.c_int(-7)
    lib.parcel_dcagmres_c_(I(0), P(x), P(b), I(n), I(n), I(nnz), I(1), P(val), P(rp), P(col), I(2), I(0), I(0),
                           I(4), ctypes.byref(niter), D(1e-9), P(hist), I(0), I(4), I(1), I(1), I(2000), I(4), I(2),
                           I(1), I(3), ctypes.byref(iret))
    xo, ho, no, io = orc.cagmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=400, rtol=1e-9, s=4,
                                 t=2, basis=1, qrflag=3)
    assert (niter.value, iret.value) == (no, io)
    assert np.array_equal(hist[:no], ho) and np.array_equal(x, xo)
    xh, xl, hist = np.zeros(n), np.zeros(n), np.zeros(500)
    niter, iret = ctypes.c_int(400), ctypes.c_int(-7)
    z, zv = np.zeros(n), np.zeros(nnz)
    lib.parcel_ddcg_c_(I(0), P(xh), P(xl), P(b), P(z), I(n), I(n), I(nnz), I(1), P(val), P(zv), P(rp), P(col), I(2),
                       I(1), I(0), I(4), ctypes.byref(niter), D(1e-20), P(hist), I(0), I(5), I(1), I(1), I(2000), I(2),
                       I(2), I(2), I(2), ctypes.byref(iret))
    oh, ol, ho, no, io = orc.ddcg(rp, col, val, zv, b, z, ipreflag=2, ilu_method=1, nblocks=5, prec=2, itrmax=400,
                                  rtol=1e-20)
    assert (niter.value, iret.value) == (no, io)
    assert np.array_equal(hist[:no], ho) and np.array_equal(xh, oh) and np.array_equal(xl, ol)


def test_by_reference_dia_entry_points():
    # the DIA forms (c2fortran_kernel.f90:1228-2091) against the by-value DIA
    # entry points and the oracle
    lib = ph.load()
    g = 10
    rp, col, val = ph.poisson3d_parcel(g, g, g)
    b, _ = ph.poisson3d_rhs_parcel(g, g, g)
    n = g ** 3
    off, dia = ph.crs_to_dia(rp, col, val)
    nnd = int(off.size)
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    D = lambda v: ctypes.byref(ctypes.c_double(v))
    lib.parcel_init_c_(I(0))
    common = lambda: (I(2), I(0), I(0), I(4))
    tail = lambda: (I(0), I(4), I(1), I(1), I(2000))
    x, hist = np.zeros(n), np.zeros(500)
    niter, iret = ctypes.c_int(400), ctypes.c_int(-7)
    lib.parcel_dcg_dia_c_(I(0), P(x), P(b), I(n), I(n), I(1), P(dia), P(off), I(nnd), *common(),
                          ctypes.byref(niter), D(1e-9), P(hist), *tail(), ctypes.byref(iret))
    xo, ho, no, io = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=4, itrmax=400, rtol=1e-9)
    assert (niter.value, iret.value) == (no, io)
    assert np.array_equal(hist[:no], ho) and np.array_equal(x, xo)
    for name, extra, solver, kw in (("parcel_dbicgstab_dia_c_", (), "bicgstab", {}),
                                    ("parcel_dgmres_dia_c_", (6, 1), "gmres", {}),
                                    ("parcel_dcagmres_dia_c_", (4, 2, 1, 3), "cagmres", {}),
                                    ("parcel_dcbcg_dia_c_", (4, 1, 30, 4, 10, 1, 3), "cbcg", {})):
        x, hist = np.zeros(n), np.zeros(500)
        niter, iret = ctypes.c_int(400), ctypes.c_int(-7)
        getattr(lib, name)(I(0), P(x), P(b), I(n), I(n), I(1), P(dia), P(off), I(nnd), *common(),
                           ctypes.byref(niter), D(1e-9), P(hist), *tail(), *[I(v) for v in extra],
                           ctypes.byref(iret))
        x2 = np.zeros(n)
        h2, n2, i2 = ph.solve_dia(solver, x2, b, off, dia, ipreflag=2, ILU_method=0, itrmax=400, rtolmax=1e-9,
                                  precon_thblock=4, extra=extra)
        assert (niter.value, iret.value) == (n2, i2), name
        assert np.array_equal(hist[:n2], h2) and np.array_equal(x, x2), name
    # CA-Arnoldi: one power step, every component of the eigenvector (DIA)
    ev = [np.zeros(40) for _ in range(2)] + [np.zeros(40 * n), np.zeros(40)]
    mx, mn, iret = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(-7)
    x0 = np.random.default_rng(1).standard_normal(n)
    lib.parcel_dcaarnoldi_dia_c_(I(0), P(x0), P(b), I(n), I(n), I(1), P(dia), P(off), I(nnd), *common(), I(0),
                                 I(4), I(1), I(2000), I(4), I(10), I(1), I(3), P(ev[0]), P(ev[1]), P(ev[2]),
                                 P(ev[3]), ctypes.byref(mx), ctypes.byref(mn), ctypes.byref(iret))
    e, en, vec = orc.power_eigen(rp, col, val, x0, ipreflag=2, ilu_method=0, nblocks=4, itrmax=1)
    assert (mx.value, mn.value, iret.value) == (1, 1, 0)
    assert ev[0][0] == e and ev[1][0] == 0.0 and ev[3][0] == en
    assert np.array_equal(ev[2][:n], vec)
