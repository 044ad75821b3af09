"""Learned-factor preconditioning of random-coefficient diffusion systems.

The workload: 2D five-point diffusion on an nx*ny grid, conductivity
exp(g) with g a squared-exponential Gaussian random field (correlation
length 16 cells, sigma 2); right-hand sides b ~ N(0,1).  Training systems
use field seeds 1000..3047, test systems 3048..3111; the right-hand side of
the system with field seed k is the normal stream of seed RHS_SEED0 + k.

Training (on the device, through the library's network kernels):
  1. generate the training systems in chunks, solve each chunk with the
     batched IC(0) PCG to produce the solutions the loss needs;
  2. Adam on minibatches of the pool, loss ||Lh Lh^T xh - bh||^2 / ||bh||^2
     (Jacobi-scaled system, Lh the network's factor).
Use: `solve_learned` predicts every test system's factor with one batched
inference and runs the batched FP64 PCG with the factor applied in FP32; a
system that does not converge falls back to the plain FP64 block-Jacobi
ILU(0) PCG.

    python learned.py train --out weights/mpnn_grf512.npy [--systems 2048]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

import parcel_hip as ph  # noqa: E402

NX = NY = 512
ELL, SIGMA = 16.0, 2.0
TRAIN_SEED0, N_TRAIN = 1000, 2048
TEST_SEED0, N_TEST = 3048, 64
RHS_SEED0 = 1_000_000
WEIGHTS = os.path.join(_HERE, "weights", "mpnn_grf512.npy")

C, NL, NO = 32, 6, 8
LAYER = C * 2 * C + C


def init_params(seed=0, scale=0.05):
    """Small random layers and a zero head: the untrained factor is the unit
    lower triangle plus the scaled matrix's strictly lower part."""
    rng = np.random.default_rng(seed)
    p = np.zeros(ph.mpnn_nparams(), np.float32)
    for l in range(NL):
        p[l * LAYER: l * LAYER + 2 * C * C] = rng.standard_normal(2 * C * C) * scale
    return p


def systems(seed0, nsys, nx=NX, ny=NY):
    """(rowptr, col, vals[nsys, nnz], b[nsys, n]) for field seeds seed0 .. seed0+nsys-1."""
    rp, col, vals = ph.grf_diffusion2d(nx, ny, seed0, nsys, ell=ELL, sigma=SIGMA)
    b = np.stack([ph.normal(nx * ny, RHS_SEED0 + seed0 + s) for s in range(nsys)])
    return rp, col, vals, np.ascontiguousarray(b)


def load_weights(path=WEIGHTS):
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (ph.mpnn_nparams(),):
        raise ValueError("weights file %s does not hold %d float32 parameters" % (path, ph.mpnn_nparams()))
    return p


def solve_learned(net, rp, col, vals, b, x=None, rtol=1e-10, itrmax=20000, precision=ph.BATCH_APPLY32,
                  inner_rtol=1e-4, profile=False, fallback_blocks=32):
    """Factor prediction + batched solve (FP64 PCG with the factor applied in
    FP32 by default; precision 32: FP64 refinement around FP32 PCG, itrmax
    then counts corrections).  A system that ends with iret != 0 is solved
    again by the plain FP64 PCG + block-Jacobi ILU(0) (fallback_blocks
    blocks), which then decides its convergence.  Returns (x, niter, iret,
    stats, seconds); stats["fell_back"] lists the systems that fell back."""
    nsys, n = b.shape
    with ph.Batch(rp, col, vals, ph.BATCH_FACTOR) as B:
        B.set_inner(inner_rtol, 20000)
        B.set_profile(profile)
        B.set_fallback(ph.PRECON_BJ, ph.ILU0, fallback_blocks, 20000)
        x = np.zeros_like(b) if x is None else x
        t0 = time.perf_counter()
        net.factor(B)
        _, niter, iret = B.solve(x, b, itrmax, rtol, precision=precision, history=False)
        dt = time.perf_counter() - t0
        st = B.stats()
        st["fell_back"] = [int(s) for s in np.flatnonzero(B.fell_back())]
        return x, niter, iret, st, dt


def train(out, n_systems=N_TRAIN, chunk=64, steps=3000, minibatch=4, lr=1e-3, seed=0, log=print, nx=NX, ny=NY,
          seed0=TRAIN_SEED0):
    n = nx * ny
    p0 = init_params(seed)
    T = ph.MpnnTrainer(n, nx, n_systems, p0)
    t0 = time.perf_counter()
    its = []
    for c0 in range(0, n_systems, chunk):
        k = min(chunk, n_systems - c0)
        rp, col, vals, b = systems(seed0 + c0, k, nx, ny)
        with ph.Batch(rp, col, vals, ph.BATCH_IC0) as B:
            x = np.zeros_like(b)
            _, niter, iret = B.solve(x, b, 5000, 1e-8, history=False)
            if np.any(iret != 0):
                raise RuntimeError("training solve did not converge (chunk %d)" % c0)
            its.append(int(niter.max()))
            T.add(B, x, b)
        log("pool %d/%d systems (IC(0) PCG its max %d) %.1fs" % (c0 + k, n_systems, its[-1],
                                                                  time.perf_counter() - t0))
    rng = np.random.default_rng(seed + 1)
    curve = []
    t1 = time.perf_counter()
    for s in range(steps):
        # cosine decay to lr/20
        lr_s = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * s / steps)))
        loss = T.step(rng.choice(n_systems, minibatch, replace=False), float(lr_s))
        curve.append(loss)
        if s % 250 == 0 or s == steps - 1:
            log("step %d loss %.5f (mean of last 100 %.5f) %.1fs" % (s, loss, float(np.mean(curve[-100:])),
                                                                     time.perf_counter() - t1))
    p = T.params()
    T.close()
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    np.save(out, p)
    meta = dict(grid=[nx, ny], ell=ELL, sigma=SIGMA, train_seeds=[seed0, seed0 + n_systems - 1],
                rhs_seed0=RHS_SEED0, steps=steps, minibatch=minibatch, lr=lr, init_seed=seed,
                loss_first100=float(np.mean(curve[:100])), loss_last100=float(np.mean(curve[-100:])),
                pool_seconds=t1 - t0, train_seconds=time.perf_counter() - t1,
                loss_curve_every50=[float(v) for v in curve[::50]])
    with open(os.path.splitext(out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    return p, meta


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--systems", type=int, default=N_TRAIN)
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--minibatch", type=int, default=4)
    t.add_argument("--lr", type=float, default=1e-3)
    args = ap.parse_args()
    ph.init(0)
    if args.cmd == "train":
        _, meta = train(args.out, args.systems, steps=args.steps, minibatch=args.minibatch, lr=args.lr,
                        log=lambda *a: print(*a, flush=True))
        print(json.dumps({k: v for k, v in meta.items() if k != "loss_curve_every50"}), flush=True)


if __name__ == "__main__":
    main()
