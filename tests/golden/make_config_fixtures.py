"""Writes tests/golden/configs.json: the oracle's results on the full-size
benchmark configurations, for the GPU tests to compare against bit for bit
without re-running multi-minute CPU solves on the GPU host.

  C2  3D upwind convection-diffusion 128^3 (cell Peclet 0.5, velocity
      direction (1, 0.5, 0.25)), x_true ~ U[-1,1] seed 1, b = A x_true (FP64,
      CRS order), BJ ILU(0) with 64 blocks, rtol 1e-9: GMRES(30) with MGS;
      CA-GMRES(s=8, t=4) with the Newton basis (Leja-ordered Ritz values of the
      32 x 32 Hessenberg) in the reference's form (basis 1) and the
      consistent form (basis 2), TSQR.
  C3  3D harmonic-mean diffusion 96^3, kappa log-uniform in [1e-4, 1e4]
      (seed 2), x_true ~ U[-1,1] seed 102, b = A x_true in double-double,
      dd PCG with BJ ILU(0) (32 blocks) to rtol 1e-20.
  C4  3D 7-point Poisson 256^3, x_true ~ U[-1,1] seed 3, b = A x_true:
      FP64 PCG + BJ ILU(0) (256 blocks) to 1e-10, and the FP64 refinement
      around FP32 PCG + FP32 BJ ILU(0) (inner rtol 1e-4) to 1e-10, whose
      blocks (the grid planes) make the FP32 apply the line-scan form of the
      block IC(0) factor (= block ILU(0) of the symmetric matrix in exact
      arithmetic).

Each entry holds the residual history (exact doubles), niter, iret and the
SHA-256 of the solution's bytes (little-endian float64).  The CPU tests
re-run a prefix of every entry and check it against this file.

    python tests/golden/make_config_fixtures.py [--only c2,c3,c4]   (c4mixed: the mixed entry only)
"""
import argparse
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "parcel-legacy_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

OUT = os.path.join(HERE, "configs.json")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def c2_problem(g=128):
    import oracle as orc
    import parcel_hip as ph
    rp, col, val = ph.convdiff3d(g, g, g, 0.5, (1.0, 0.5, 0.25))
    b = orc.spmv(rp, col, val, ph.uniform(g ** 3, 1))
    return rp, col, val, b


def c3_problem(g=96):
    import parcel_hip as ph
    n = g ** 3
    rp, col, val = ph.diffusion3d(g, g, g, ph.loguniform(n, 2, 1e-4, 1e4))
    bh, bl = ph.spmv_dd(rp, col, val, ph.uniform(n, 102))
    return rp, col, val, bh, bl


def c4_problem(g=256):
    import oracle as orc
    import parcel_hip as ph
    rp, col, val = ph.poisson3d_spd(g, g, g)
    b = orc.spmv(rp, col, val, ph.uniform(g ** 3, 3))
    return rp, col, val, b


# solver settings of every entry (the tests read them from here too)
C2_GMRES = dict(m=30, gsflag=1, nblocks=64, rtol=1e-9, itrmax=5000)
C2_CAGMRES = dict(s=8, t=4, qrflag=3, nblocks=64, rtol=1e-9, itrmax=5000)
C3_DDCG = dict(nblocks=32, rtol=1e-20, itrmax=20000)
C4_BJ = dict(nblocks=256, rtol=1e-10, itrmax=20000)
C4_MIXED = dict(nblocks=256, rtol=1e-10, itrmax=50, inner_rtol=1e-4, inner_itrmax=20000)


def entry(hist, niter, iret, **sums):
    e = {"hist": [float(v) for v in hist], "niter": int(niter), "iret": int(iret)}
    e.update(sums)
    return e


def make_c2(out, log):
    import oracle as orc
    rp, col, val, b = c2_problem()
    t = time.time()
    c = C2_GMRES
    x, h, ni, ir = orc.gmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=c["nblocks"], itrmax=c["itrmax"],
                             rtol=c["rtol"], m=c["m"], gsflag=c["gsflag"])
    out["c2_gmres30"] = entry(h, ni, ir, x_sha256=sha(x))
    log("c2 gmres30: %d its %.1fs" % (ni, time.time() - t))
    c = C2_CAGMRES
    for basis in (1, 2):
        t = time.time()
        x, h, ni, ir = orc.cagmres(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=c["nblocks"],
                                   itrmax=c["itrmax"], rtol=c["rtol"], s=c["s"], t=c["t"], basis=basis,
                                   qrflag=c["qrflag"])
        res = float(np.linalg.norm(b - orc.spmv(rp, col, val, x)) / np.linalg.norm(b))
        out["c2_cagmres_s8t4_basis%d" % basis] = entry(h, ni, ir, x_sha256=sha(x), true_rel_res=res)
        log("c2 cagmres basis %d: %d its, true res %.3e %.1fs" % (basis, ni, res, time.time() - t))


def make_c3(out, log):
    import oracle as orc
    rp, col, val, bh, bl = c3_problem()
    c = C3_DDCG
    t = time.time()
    xh, xl, h, ni, ir = orc.ddcg(rp, col, val, None, bh, bl, ipreflag=2, ilu_method=0, nblocks=c["nblocks"], prec=2,
                                 itrmax=c["itrmax"], rtol=c["rtol"])
    out["c3_ddcg"] = entry(h, ni, ir, xh_sha256=sha(xh), xl_sha256=sha(xl))
    log("c3 ddcg: %d its %.1fs" % (ni, time.time() - t))


def make_c4(out, log):
    import oracle as orc
    rp, col, val, b = c4_problem()
    c = C4_BJ
    t = time.time()
    x, h, ni, ir = orc.pcg(rp, col, val, b, ipreflag=2, ilu_method=0, nblocks=c["nblocks"], itrmax=c["itrmax"],
                           rtol=c["rtol"])
    out["c4_bj"] = entry(h, ni, ir, x_sha256=sha(x))
    log("c4 bj: %d its %.1fs" % (ni, time.time() - t))
    make_c4_mixed(out, log, (rp, col, val, b))


def c4_mixed_precon(rp, col, val, g=256):
    """The FP32 inner preconditioner of the C4 mixed solve: the 256 blocks
    are the grid planes, so block ILU(0) of the symmetric matrix is applied
    as line scans of the block IC(0) factor."""
    import oracle as orc
    nb = C4_MIXED["nblocks"]
    F64, F32 = orc.derive(*orc.ic0_grid(rp, col, val, g, nblocks=nb), g)
    return orc.FactorPrecon(F64, F32, g, nstrip=nb)


def make_c4_mixed(out, log, prob=None):
    import oracle as orc
    rp, col, val, b = prob if prob is not None else c4_problem()
    c = C4_MIXED
    t = time.time()
    M = c4_mixed_precon(rp, col, val)
    x, h, ni, ir, inner = orc.pcg_mixed_factor(rp, col, val, b, M, itrmax=c["itrmax"], rtol=c["rtol"],
                                               inner_itrmax=c["inner_itrmax"], inner_rtol=c["inner_rtol"])
    out["c4_mixed_bj"] = entry(h, ni, ir, x_sha256=sha(x), inner_total=int(inner))
    log("c4 mixed bj: %d outer, %d inner its %.1fs" % (ni, inner, time.time() - t))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="c2,c3,c4")
    args = ap.parse_args()
    out = json.load(open(OUT)) if os.path.exists(OUT) else {}
    log = lambda *a: print(*a, flush=True)
    for k in args.only.split(","):
        {"c2": make_c2, "c3": make_c3, "c4": make_c4, "c4mixed": make_c4_mixed}[k](out, log)
        with open(OUT, "w") as f:
            json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
