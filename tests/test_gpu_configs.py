"""The benchmark configurations at their full sizes on the GPU, through the C
ABI, bit for bit against the oracle.

C2-C4 compare with tests/golden/configs.json, the oracle's histories, niter,
iret and solution hashes on the same inputs (written by
tests/golden/make_config_fixtures.py; tests/test_configs_fixture.py re-runs a
prefix of each on the CPU), and check what the size does not change: the
true residual of the returned x.  C5 runs the trained network on the 64 test
systems and compares two of them with the oracle PCG given the GPU's factor.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle as orc
import parcel_hip as ph

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_config_fixtures as F  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(HERE, "golden", "configs.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def same(g, hist, niter, iret):
    assert (niter, iret) == (g["niter"], g["iret"])
    assert np.array_equal(hist, np.array(g["hist"]))


@pytest.fixture(scope="module")
def c2():
    say("c2: 128^3 convection-diffusion")
    return F.c2_problem()


def true_res(rp, col, val, x, b):
    return float(np.linalg.norm(b - orc.spmv(rp, col, val, x)) / np.linalg.norm(b))


def test_c2_gmres30(c2):
    rp, col, val, b = c2
    c = F.C2_GMRES
    say("c2 gmres(30) bj-ilu0 64 blocks")
    x = np.zeros_like(b)
    hist, niter, iret = ph.dgmres(x, b, rp, col, val, GMRES_m=c["m"], GMRES_GSflag=c["gsflag"], ipreflag=2,
                                  ILU_method=0, precon_thblock=c["nblocks"], itrmax=c["itrmax"], rtolmax=c["rtol"])
    g = GOLD["c2_gmres30"]
    same(g, hist, niter, iret)
    assert sha(x) == g["x_sha256"]
    assert iret == 0 and true_res(rp, col, val, x, b) < c["rtol"]


@pytest.mark.parametrize("basis", [1, 2])
def test_c2_cagmres_newton(c2, basis):
    rp, col, val, b = c2
    c = F.C2_CAGMRES
    say("c2 ca-gmres(8,4) newton basis", basis)
    x = np.zeros_like(b)
    hist, niter, iret = ph.dcagmres(x, b, rp, col, val, CAGMRES_sstep=c["s"], CAGMRES_tstep=c["t"],
                                    CAGMRES_basis=basis, CAGMRES_QRflag=c["qrflag"], ipreflag=2, ILU_method=0,
                                    precon_thblock=c["nblocks"], itrmax=c["itrmax"], rtolmax=c["rtol"])
    g = GOLD["c2_cagmres_s8t4_basis%d" % basis]
    same(g, hist, niter, iret)
    assert sha(x) == g["x_sha256"]
    # the same x (hash above); the norms of numpy's BLAS may round differently by host
    res = true_res(rp, col, val, x, b)
    assert abs(res - g["true_rel_res"]) <= 1e-10 * g["true_rel_res"]
    if basis == 2:
        # the consistent conjugate-pair step: the estimate is the true residual
        assert res < c["rtol"]
    else:
        # the reference's step with complex Ritz shifts: the estimate drifts
        assert res > 100 * hist[-1]


def test_c3_ddcg():
    say("c3: 96^3 log-uniform diffusion, dd-PCG bj-ilu0 32 blocks to 1e-20")
    rp, col, val, bh, bl = F.c3_problem()
    c = F.C3_DDCG
    n = bh.size
    xh, xl = np.zeros(n), np.zeros(n)
    hist, niter, iret = ph.ddcg(xh, xl, bh, bl, rp, col, val, None, ipreflag=2, ILU_method=0,
                                precon_thblock=c["nblocks"], itrmax=c["itrmax"], rtolmax=c["rtol"], PRECISION_A=1,
                                PRECISION_b=2, PRECISION_x=2, PRECISION_PRECONDITION=2)
    g = GOLD["c3_ddcg"]
    same(g, hist, niter, iret)
    assert sha(xh) == g["xh_sha256"] and sha(xl) == g["xl_sha256"]
    assert iret == 0 and hist[-1] < 1e-19


@pytest.fixture(scope="module")
def c4():
    say("c4: 256^3 poisson")
    return F.c4_problem()


def test_c4_bj(c4):
    rp, col, val, b = c4
    c = F.C4_BJ
    say("c4 fp64 pcg bj-ilu0 256 blocks")
    x = np.zeros_like(b)
    hist, niter, iret = ph.dcg(x, b, rp, col, val, ipreflag=2, ILU_method=0, precon_thblock=c["nblocks"],
                               itrmax=c["itrmax"], rtolmax=c["rtol"])
    g = GOLD["c4_bj"]
    same(g, hist, niter, iret)
    assert sha(x) == g["x_sha256"]
    assert true_res(rp, col, val, x, b) < 2 * c["rtol"]


def test_c4_mixed_bj(c4):
    rp, col, val, b = c4
    c = F.C4_MIXED
    say("c4 fp64 refinement around fp32 pcg + fp32 bj-ilu0 256 blocks")
    p = ph.default_param(solver=11, ipreflag=2, ILU_method=0, precon_thblock=c["nblocks"])
    with ph.Plan(rp, col, val, param=p) as plan:
        plan.set_inner(c["inner_rtol"], c["inner_itrmax"])
        x = np.zeros_like(b)
        hist, niter, iret = plan.solve(x, b, c["itrmax"], c["rtol"])
        st = plan.stats()
    # the blocks are the grid planes: the FP32 apply is the line-scan form of
    # the block IC(0) factor (block ILU(0) of the symmetric matrix)
    assert st["bj_lscan"] == 1
    g = GOLD["c4_mixed_bj"]
    same(g, hist, niter, iret)
    assert st["inner_niter"] == g["inner_total"]
    assert sha(x) == g["x_sha256"]
    assert iret == 0 and true_res(rp, col, val, x, b) < c["rtol"]


def test_c5_learned_64_systems():
    import learned as L
    say("c5: trained network on the 64 test systems, batched fp32-apply pcg")
    rp, col, vals, b = L.systems(L.TEST_SEED0, L.N_TEST)
    nsys, n = b.shape
    p = L.load_weights()
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        net = ph.Mpnn(p)
        net.factor(B)
        net.close()
        Lb = B.factor()
        x = np.zeros_like(b)
        hist, niter, iret = B.solve(x, b, 20000, 1e-10, precision=ph.BATCH_APPLY32)
        y = np.zeros_like(b)
        B.spmv(x, y)
    assert np.all(iret == 0)
    res = np.linalg.norm(b - y, axis=1) / np.linalg.norm(b, axis=1)
    say("c5 true residuals max %.3e, iterations %d..%d" % (res.max(), niter.min(), niter.max()))
    assert res.max() < 1e-10
    for s in (0, nsys - 1):
        sl = slice(s * n, (s + 1) * n)
        F64, F32 = orc.derive(Lb[0][sl], Lb[1][sl], Lb[2][sl], L.NX)
        M = orc.FactorPrecon(F64, F32, L.NX, apply32=True)
        xo, ho, no, io = orc.pcg_factor(rp, col, vals[s], b[s], M, itrmax=20000, rtol=1e-10)
        assert (niter[s], iret[s]) == (no, io)
        assert np.array_equal(hist[s, :no], ho) and np.array_equal(x[s], xo)
