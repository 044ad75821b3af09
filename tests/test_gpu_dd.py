"""GPU parity tests of the double-double PCG (src/ddpcg.f90:247-563)
through the C ABI: history, x (hi and lo), niter and iret bit-exact against
the oracle, whose dd operations, SpMV/sweep orders and tiled dd reductions
the kernels reproduce; plus the reference's dd golden history."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


PRECONS = {
    "none": dict(ipreflag=0),
    "pj": dict(ipreflag=1),
    "ilu0_4": dict(ipreflag=2, ILU_method=0, precon_thblock=4),
    "ilud_6": dict(ipreflag=2, ILU_method=1, precon_thblock=6),
    "mlud_3": dict(ipreflag=2, ILU_method=2, precon_thblock=3),
}


def _run(rp, col, vh, vl, bh, bl, kw, pa, pb, pp, itrmax, rtol):
    n = rp.shape[0] - 1
    xh, xl = np.zeros(n), np.zeros(n)
    hist, niter, iret = ph.ddcg(xh, xl, bh, bl, rp, col, vh, vl, ipreflag=kw["ipreflag"],
                                ILU_method=kw.get("ILU_method", 1), precon_thblock=kw.get("precon_thblock", 1),
                                itrmax=itrmax, rtolmax=rtol, PRECISION_A=pa, PRECISION_b=pb, PRECISION_x=2,
                                PRECISION_PRECONDITION=pp)
    oh, ol, ho, no, io = orc.ddcg(rp, col, vh, vl if pa == 2 else None, bh, bl if pb == 2 else None,
                                  ipreflag=kw["ipreflag"], ilu_method=kw.get("ILU_method", 1),
                                  nblocks=kw.get("precon_thblock", 1), prec=pp, itrmax=itrmax, rtol=rtol)
    return (xh, xl, hist, niter, iret), (oh, ol, ho, no, io)


@pytest.mark.parametrize("pp", [1, 2])
@pytest.mark.parametrize("pname", sorted(PRECONS))
def test_ddcg_bit_exact(pname, pp):
    say("ddcg", pname, pp)
    rp, col, val = ph.poisson3d_parcel(16, 16, 16)
    b, _ = ph.poisson3d_rhs_parcel(16, 16, 16)
    vlo = val * 2.0 ** -58
    blo = b * 2.0 ** -57
    for pa, pb in ((1, 1), (2, 2)):
        (xh, xl, h, ni, ir), (oh, ol, ho, no, io) = _run(rp, col, val, vlo, b, blo, PRECONS[pname], pa, pb, pp,
                                                         1500, 1e-22)
        assert (ni, ir) == (no, io)
        assert np.array_equal(h, ho)
        assert np.array_equal(xh, oh) and np.array_equal(xl, ol)


def test_ddcg_capped_tiling_and_itrmax():
    say("ddcg 2d 760")
    rp, col, val = ph.poisson2d(760, 760, -1.0)
    b = ph.uniform(rp.shape[0] - 1, 51)
    for itrmax in (0, 1, 7):
        (xh, xl, h, ni, ir), (oh, ol, ho, no, io) = _run(rp, col, val, None, b, None, PRECONS["ilu0_4"], 1, 1, 2,
                                                         itrmax, 1e-30)
        assert (ni, ir) == (no, io)
        assert np.array_equal(h, ho) and np.array_equal(xh, oh) and np.array_equal(xl, ol)


def test_ddcg_matches_reference_dd_template():
    say("ddcg golden 160")
    gold = np.array(json.load(open(os.path.join(HERE, "golden", "templates.json")))
                    ["template_cpu/STDOUT_run_dd"]["histories"][0])
    rp, col, val = ph.poisson3d_parcel(160, 160, 160)
    b, _ = ph.poisson3d_rhs_parcel(160, 160, 160)
    n = b.size
    xh, xl = np.zeros(n), np.zeros(n)
    hist, niter, iret = ph.ddcg(xh, xl, b, np.zeros(n), rp, col, val, np.zeros_like(val), ipreflag=2, ILU_method=1,
                                precon_thblock=80, itrmax=2000, rtolmax=1e-9)
    assert niter == gold.shape[0] == 291 and iret == 0
    assert np.max(np.abs(hist - gold[:, 1]) / gold[:, 1]) <= 1e-12
