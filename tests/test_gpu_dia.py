"""GPU tests of the DIA-format entry points (parcel_dcg_dia and friends,
src/wrapper.f90:454-1558): on every solver the DIA form equals the CRS form
bit for bit (the fused CG kernel reads the diagonals; dia2crs feeds the
preconditioner set-up), and the DIA columns of the reference's own goldens
(template_cpu / template_gpu STDOUT_run: 160^3, CRS and DIA printed side by
side) are reproduced to 1e-12."""
import json
import os

import numpy as np
import pytest

import parcel_hip as ph
from problems import convdiff3d

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


PRECONS = {"none": dict(ipreflag=0), "pj": dict(ipreflag=1), "ilu0_4": dict(ipreflag=2, ILU_method=0,
                                                                           precon_thblock=4),
           "ilud_8": dict(ipreflag=2, ILU_method=1, precon_thblock=8)}


def crs_solve(solver, x, b, rp, col, val, kw, extra):
    if solver == "cg":
        return ph.dcg(x, b, rp, col, val, **kw)
    if solver == "bicgstab":
        return ph.dbicgstab(x, b, rp, col, val, **kw)
    if solver == "gmres":
        return ph.dgmres(x, b, rp, col, val, GMRES_m=extra[0], GMRES_GSflag=extra[1], **kw)
    if solver == "cagmres":
        return ph.dcagmres(x, b, rp, col, val, CAGMRES_sstep=extra[0], CAGMRES_tstep=extra[1],
                           CAGMRES_basis=extra[2], CAGMRES_QRflag=extra[3], **kw)
    return ph.dcbcg(x, b, rp, col, val, CBCG_kstep=extra[0], CBCG_Eigenflag=extra[1], power_method_itrmax=extra[2],
                    CAARNOLDI_sstep=extra[3], CAARNOLDI_tstep=extra[4], CAARNOLDI_basis=extra[5],
                    CAARNOLDI_QRflag=extra[6], **kw)


SOLVERS = {"cg": (), "bicgstab": (), "gmres": (8, 1), "cagmres": (4, 2, 2, 5), "cbcg": (4, 1, 100, 4, 10, 1, 3)}


@pytest.mark.parametrize("pname", sorted(PRECONS))
@pytest.mark.parametrize("solver", sorted(SOLVERS))
@pytest.mark.parametrize("case", ["parcel3d_14", "convdiff_10"])
def test_dia_equals_crs(case, solver, pname):
    print("dia", case, solver, pname, flush=True)
    if case == "parcel3d_14":
        rp, col, val = ph.poisson3d_parcel(14, 14, 14)
    else:
        if solver in ("cg", "cbcg"):
            pytest.skip("CG and CBCG need a symmetric matrix")
        rp, col, val = convdiff3d(10, peclet=1.0)[:3]
    off, dia = ph.crs_to_dia(rp, col, val)
    n = rp.shape[0] - 1
    b = ph.uniform(n, 13)
    kw = dict(PRECONS[pname], itrmax=600, rtolmax=1e-9)
    x1, x2 = np.zeros(n), np.zeros(n)
    h1, n1, i1 = crs_solve(solver, x1, b, rp, col, val, kw, SOLVERS[solver])
    h2, n2, i2 = ph.solve_dia(solver, x2, b, off, dia, extra=SOLVERS[solver], **kw)
    assert (n1, i1) == (n2, i2)
    assert np.array_equal(h1, h2) and np.array_equal(x1, x2)


@pytest.mark.parametrize("ilu,nblk,gname", [(1, 80, "template_cpu/STDOUT_run"), (0, 4, "template_gpu/STDOUT_run")])
def test_dia_reference_golden(ilu, nblk, gname):
    print("dia golden", gname, flush=True)
    gold = np.array(json.load(open(os.path.join(HERE, "golden", "templates.json")))[gname]["histories"][0])
    rp, col, val = ph.poisson3d_parcel(160, 160, 160)
    b, _ = ph.poisson3d_rhs_parcel(160, 160, 160)
    off, dia = ph.crs_to_dia(rp, col, val)
    x = np.zeros(b.size)
    hist, niter, iret = ph.solve_dia("cg", x, b, off, dia, ipreflag=2, ILU_method=ilu, precon_thblock=nblk,
                                     itrmax=2000, rtolmax=1e-9)
    assert iret == 0 and niter == gold.shape[0]
    rel = np.abs(hist - gold[:, 2]) / gold[:, 2]
    print("max rel dev vs the DIA column", rel.max(), flush=True)
    assert rel.max() <= 1e-12
