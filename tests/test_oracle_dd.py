"""CPU tests of the oracle's double-double path (src/dd_func.f90,
src/ddpcg.f90, src/ddmatrix.f90).

* error-free transforms: dd_add / dd_mul of two doubles are exact (checked
  in exact rational arithmetic), and the dd operations on dd operands are
  accurate to about 2^-104;
* dd dot products and SpMV against exact rational sums;
* the reference golden template_cpu/STDOUT_run_dd (dd PCG, D-ILU in dd, 80
  blocks, 160^3, 291 residuals) within 1e-12 (slow);
* dd PCG against FP64 PCG on small systems (equal to FP64 round-off) and
  reaching a residual far below FP64 resolution."""
import json
import os
from fractions import Fraction as F

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph

HERE = os.path.dirname(os.path.abspath(__file__))


def _rand(n, seed, spread=20):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) * np.exp(rng.uniform(-spread, spread, n))


def _dd(n, seed):
    h = _rand(n, seed)
    l = h * 2.0 ** -60 * np.random.default_rng(seed + 1).uniform(-1, 1, n)
    # normalise so that |lo| <= ulp(hi)/2
    s = h + l
    return s, l - (s - h)


def test_two_sum_and_two_product_are_exact():
    a, b = _rand(2000, 1), _rand(2000, 2)
    z = np.zeros_like(a)
    h, l = orc.dd_op("add", a, z, b, z)
    assert all(F(h[i]) + F(l[i]) == F(a[i]) + F(b[i]) for i in range(a.size))
    h, l = orc.dd_op("mul", a, z, b, z)
    assert all(F(h[i]) + F(l[i]) == F(a[i]) * F(b[i]) for i in range(a.size))
    h, l = orc.dd_op("sub", a, z, b, z)
    assert all(F(h[i]) + F(l[i]) == F(a[i]) - F(b[i]) for i in range(a.size))


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_dd_operations_accuracy(op):
    ah, al = _dd(500, 3)
    bh, bl = _dd(500, 4)
    h, l = orc.dd_op(op, ah, al, bh, bl)
    worst = 0.0
    for i in range(ah.size):
        a, b = F(ah[i]) + F(al[i]), F(bh[i]) + F(bl[i])
        exact = {"add": a + b, "sub": a - b, "mul": a * b, "div": a / b}[op]
        got = F(h[i]) + F(l[i])
        if exact != 0:
            worst = max(worst, abs(float((got - exact) / exact)))
    bound = 2.0 ** -100 if op in ("mul", "div") else 2.0 ** -102
    if op in ("add", "sub"):
        # cancellation: relative to the larger operand
        worst = 0.0
        for i in range(ah.size):
            a, b = F(ah[i]) + F(al[i]), F(bh[i]) + F(bl[i])
            exact = a + b if op == "add" else a - b
            got = F(h[i]) + F(l[i])
            worst = max(worst, abs(float((got - exact) / max(abs(a), abs(b)))))
    assert worst <= bound


def test_dd_sqrt_accuracy_and_edges():
    ah, al = _dd(300, 5)
    ah, al = np.abs(ah), np.where(ah < 0, -al, al)
    h, l = orc.dd_op("sqrt", ah, al)
    for i in range(ah.size):
        a = F(ah[i]) + F(al[i])
        r = F(h[i]) + F(l[i])
        assert abs(float((r * r - a) / a)) <= 2.0 ** -100
    # a negative argument returns 0; a zero argument yields NaN (no early return in dd_sqrt)
    h, l = orc.dd_op("sqrt", np.array([-4.0, 0.0]), np.zeros(2))
    assert h[0] == 0.0 and np.isnan(h[1])


def test_dd_dot_and_spmv_against_exact_sums():
    n = 3000
    xh, xl = _dd(n, 6)
    yh, yl = _dd(n, 7)
    exact = sum((F(xh[i]) + F(xl[i])) * (F(yh[i]) + F(yl[i])) for i in range(n))
    scale = sum(abs((F(xh[i]) + F(xl[i])) * (F(yh[i]) + F(yl[i]))) for i in range(n))
    for red in (orc.TILED, orc.SEQ):
        rh, rl = orc.dd_dot(xh, xl, yh, yl, red)
        assert abs(float((F(rh) + F(rl) - exact) / scale)) <= 4 * n * 2.0 ** -104
    rp, col, val = ph.poisson3d_parcel(6, 6, 6)
    m = 216
    vlo = val * 2.0 ** -70
    xh, xl = _dd(m, 8)
    yh, yl = orc.dd_spmv(rp, col, val, vlo, xh, xl)
    for k in range(m):
        ex = sum((F(val[j]) + F(vlo[j])) * (F(xh[col[j] - 1]) + F(xl[col[j] - 1])) for j in range(rp[k] - 1, rp[k + 1] - 1))
        sc = sum(abs((F(val[j]) + F(vlo[j])) * (F(xh[col[j] - 1]) + F(xl[col[j] - 1]))) for j in range(rp[k] - 1, rp[k + 1] - 1))
        assert abs(float((F(yh[k]) + F(yl[k]) - ex) / sc)) <= 16 * 2.0 ** -104


@pytest.mark.parametrize("kw", [dict(ipreflag=0), dict(ipreflag=1), dict(ipreflag=2, ilu_method=1, nblocks=4),
                                dict(ipreflag=2, ilu_method=2, nblocks=3), dict(ipreflag=2, ilu_method=0, nblocks=4)],
                         ids=["none", "pj", "ilud", "mlud", "ilu0"])
def test_ddcg_tracks_fp64_and_goes_below_it(kw):
    rp, col, val = ph.poisson3d_parcel(12, 12, 12)
    b, _ = ph.poisson3d_rhs_parcel(12, 12, 12)
    _, _, hd, nd, _ = orc.ddcg(rp, col, val, None, b, None, rtol=1e-9, itrmax=2000, **kw)
    _, h64, n64, _ = orc.pcg(rp, col, val, b, rtol=1e-9, itrmax=2000, ipreflag=kw["ipreflag"],
                             ilu_method=kw.get("ilu_method", 1), nblocks=kw.get("nblocks", 1))
    assert nd == n64
    assert np.max(np.abs(hd - h64) / h64) < 1e-12
    # far below FP64: the dd solve continues to 1e-24 and its x has a dd-accurate residual
    xh, xl, hd2, nd2, ir2 = orc.ddcg(rp, col, val, None, b, None, rtol=1e-24, itrmax=3000, **kw)
    assert ir2 == 0 and hd2[-1] >= 1e-24
    rh, rl = orc.dd_spmv(rp, col, val, None, xh, xl)
    n = b.size
    res = [F(b[i]) - (F(rh[i]) + F(rl[i])) for i in range(n)]
    rn = sum(r * r for r in res)
    bn = sum(F(v) * F(v) for v in b)
    assert float(rn / bn) ** 0.5 < 1e-22


def test_dd_precondition_precision_one_uses_hi_parts_only():
    rp, col, val = ph.poisson3d_parcel(8, 8, 8)
    vlo = val * 2.0 ** -60
    rh, rl = _dd(512, 9)
    for kw in (dict(ipreflag=1), dict(ipreflag=2, ilu_method=1, nblocks=2), dict(ipreflag=2, ilu_method=0, nblocks=2)):
        s1 = orc.DDPrecon(rp, col, val, vlo, prec=1, **kw).apply(rh, rl)
        s2 = orc.DDPrecon(rp, col, val, vlo, prec=2, **kw).apply(rh, rl)
        assert not np.array_equal(s1[1], s2[1])
        # with hi-only factors, an A without lo part gives the same result at both precisions
        t1 = orc.DDPrecon(rp, col, val, None, prec=1, **kw).apply(rh, rl)
        t2 = orc.DDPrecon(rp, col, val, None, prec=2, **kw).apply(rh, rl)
        # (the dd factors keep lo parts even for an FP64 matrix, so the two agree to FP64 round-off)
        assert np.max(np.abs(t1[0] - t2[0]) / np.abs(t2[0])) < 1e-14


@pytest.mark.slow
def test_ddcg_matches_reference_dd_template():
    """template_cpu/STDOUT_run_dd: dd PCG, D-ILU (ILUD) in dd, 4 ranks x 20
    threads = 80 blocks, 160^3 PARCEL Poisson, rtol 1e-9, 291 iterations."""
    gold = np.array(json.load(open(os.path.join(HERE, "golden", "templates.json")))
                    ["template_cpu/STDOUT_run_dd"]["histories"][0])
    rp, col, val = ph.poisson3d_parcel(160, 160, 160)
    b, _ = ph.poisson3d_rhs_parcel(160, 160, 160)
    _, _, hd, nd, ird = orc.ddcg(rp, col, val, None, b, None, rtol=1e-9, itrmax=2000, ipreflag=2, ilu_method=1,
                                 nblocks=80, prec=2)
    assert nd == gold.shape[0] == 291 and ird == 0
    assert np.max(np.abs(hd - gold[:, 1]) / gold[:, 1]) <= 1e-12
