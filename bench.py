#!/usr/bin/env python3
"""Benchmark: time-to-solution of preconditioned CG at relative residual 1e-10
on one MI355X, with HBM GB/s against the roofline.

Workload (default): 3D 7-point constant-coefficient Poisson 256^3
(16,777,216 unknowns, 117,047,296 nonzeros, diag 6 / off -1), x_true ~ U[-1,1]
from SplitMix64 seed 3, b = A x_true, x0 = 0, rtol 1e-10 (FP64 residual).
Methods:
  bj        FP64 PCG, block-Jacobi ILU(0), 256 contiguous row blocks
  pj        FP64 PCG, point Jacobi (the SpMV/CG path)
  mixed_bj  FP64 iterative refinement around FP32 PCG with FP32 block-Jacobi
            ILU(0) (256 blocks), inner tolerance 1e-4
  mixed_pj  the same with FP32 point Jacobi
A "step" is one full solve from x0 = 0 with the matrix, b and x already in
HBM; `value` is the mean time-to-solution of the fastest method, and the
config reports its speed-up over the fastest plain FP64 PCG.  The timed
solves run without per-kernel events; one more solve with HIP events around
every kernel gives the per-kernel means behind `roofline`.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "parcel-legacy_amd"), os.path.join(ROOT, "oracle")]

PEAK_HBM_GBS = 8000.0
METRIC = "time-to-solution at rel res 1e-10 (s) and HBM GB/s: learned/mixed vs plain PCG"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def k1_bytes(n, nnz):
    """Fused CG kernel (SpMV + p/x update + <p,q>): CSR 12 B/nnz + row pointers
    4(n+1) + x read/write 16n + p_old 8n + s 8n + p_new 8n + q 8n."""
    return 12 * nnz + 4 * (n + 1) + 48 * n


def ilu0_apply_bytes(n, nnz_l, nnz_u):
    """Block-Jacobi ILU(0) apply (forward + backward sweep): 12 B per factor
    entry, row ids 4n per sweep, r 8n, s written 8n (forward), s read+write
    16n and u_ii 8n (backward)."""
    return 12 * (nnz_l + nnz_u) + 8 * n + 8 * n + 8 * n + 16 * n + 8 * n


def k1_bytes32(n, nnz):
    """FP32 fused CG kernel: values+columns 8 B/nnz, row pointers 4(n+1),
    x read/write 8n, p_old 4n, s 4n, p_new 4n, q 4n."""
    return 8 * nnz + 4 * (n + 1) + 24 * n


def ilu0_apply_bytes32(n, nnz_l, nnz_u):
    """FP32 block-Jacobi ILU(0) apply: 8 B per factor entry, row ids 4n per
    sweep, r 4n, s written 4n (forward), s read+write 8n and u_ii 4n."""
    return 8 * (nnz_l + nnz_u) + 8 * n + 4 * n + 4 * n + 8 * n + 4 * n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--methods", default="bj,pj,mixed_bj,mixed_pj")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--calib", action="store_true",
                    help="also run one plain 2 GiB device-to-device copy (reference for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one GPU holds one rank's whole subdomain); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import parcel_hip as ph

    ph.init(0)
    g = args.grid
    n = g ** 3
    t0 = time.perf_counter()
    rp, col, val = ph.poisson3d_spd(g, g, g)
    nnz = int(col.shape[0])
    xt = ph.uniform(n, 3)
    log("generated %d^3: n=%d nnz=%d in %.2fs" % (g, n, nnz, time.perf_counter() - t0))
    D = ph.DeviceArray
    xt_d = D(xt)
    b_d = D(n=n)
    x_d = D(n=n)
    r_d = D(n=n)

    if args.calib:
        src = D(n=1 << 28)
        dst = D(n=1 << 28)
        ph._check(ph.load().parcel_hip_memcpy(dst.ptr, src.ptr, src.nbytes, 3))
        ph.load().parcel_hip_device_synchronize()
        log("calibration copy: %d bytes read, %d bytes written" % (src.nbytes, src.nbytes))
        src.free()
        dst.free()

    methods = {}
    for m in args.methods.split(","):
        mixed = m.startswith("mixed_")
        base = m[len("mixed_"):] if mixed else m
        if base == "bj":
            kw = dict(ipreflag=2, ILU_method=0, precon_thblock=256)
        elif base == "pj":
            kw = dict(ipreflag=1)
        else:
            raise SystemExit("unknown method " + m)
        if mixed:
            kw["solver"] = 11
        itrmax = 50 if mixed else 20000
        t0 = time.perf_counter()
        plan = ph.Plan(rp, col, val, profile=0, **kw)
        if mixed:
            plan.set_inner(1e-4, 20000)
        setup = time.perf_counter() - t0
        plan.spmv(xt_d, b_d)
        for _ in range(args.warmup):
            x_d.zero()
            plan.solve(x_d, b_d, itrmax, 1e-10)
        times = []
        dev_ms = []
        st = None
        for k in range(args.steps):
            x_d.zero()
            ph.load().parcel_hip_device_synchronize()
            t1 = time.perf_counter()
            hist, niter, iret = plan.solve(x_d, b_d, itrmax, 1e-10)
            times.append(time.perf_counter() - t1)
            st = plan.stats()
            dev_ms.append(st["solve_device_ms"])
            log("%s step %d: %.4f s, niter %d, iret %d, final %.3e" % (m, k, times[-1], niter, iret, hist[-1]))
        # kernel timing pass: the same solve with HIP events around each kernel
        plan.set_profile(True)
        x_d.zero()
        plan.solve(x_d, b_d, itrmax, 1e-10)
        st = plan.stats()
        plan.set_profile(False)
        # true residual of the last solve
        plan.spmv(x_d, r_d)
        bb = b_d.numpy()
        true_res = float(np.linalg.norm(bb - r_d.numpy()) / np.linalg.norm(bb))
        k1_ms = st["spmv_ms"] / max(st["spmv_calls"], 1)
        pre_ms = st["precon_ms"] / max(st["precon_calls"], 1) if st["precon_calls"] else 0.0
        vec_ms = st["vec_ms"] / max(st["vec_calls"], 1) if st["vec_calls"] else 0.0
        sfx = "_f32" if mixed else ""
        kb = k1_bytes32(n, nnz) if mixed else k1_bytes(n, nnz)
        kernels = {"cg_spmv_fused" + sfx: {"mean_ms": k1_ms, "calls_per_solve": st["spmv_calls"], "bytes": kb,
                                           "GBs": kb / (k1_ms * 1e-3) / 1e9}}
        if base == "bj":
            pb = (ilu0_apply_bytes32 if mixed else ilu0_apply_bytes)(n, st["bj_nnz_l"], st["bj_nnz_u"])
            kernels["bj_ilu0_apply" + sfx] = {"mean_ms": pre_ms, "calls_per_solve": st["precon_calls"],
                                              "bytes": pb, "GBs": pb / (pre_ms * 1e-3) / 1e9,
                                              "levels_fwd": st["bj_levels_fwd"], "levels_bwd": st["bj_levels_bwd"]}
        kernels["vector_reductions"] = {"mean_ms": vec_ms, "calls_per_solve": st["vec_calls"]}
        methods[m] = {"time_to_solution_s": float(np.mean(times)), "device_ms": float(np.mean(dev_ms)),
                      "niter": niter, "iret": iret, "final_recurrence_res": float(hist[-1]),
                      "inner_fp32_iters": st["inner_niter"] if mixed else None,
                      "true_rel_res": true_res, "setup_s": setup, "precon_setup_s": st["precon_setup_s"],
                      "kernels": kernels}
        plan.close()

    head = min(methods, key=lambda k: methods[k]["time_to_solution_s"])
    H = methods[head]
    plain = [k for k in methods if not k.startswith("mixed_")]
    best_plain = min(plain, key=lambda k: methods[k]["time_to_solution_s"]) if plain else None
    # dominant kernel of the headline method: the one with most time per solve
    dom_name, dom = max(((k, v) for k, v in H["kernels"].items() if "bytes" in v),
                        key=lambda kv: kv[1]["mean_ms"] * kv[1]["calls_per_solve"])
    achieved = dom["GBs"]
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pcg_c4_poisson256_traffic.json")
    if g == 256 and os.path.exists(tf):
        # HBM bytes per launch of the same kernel on the same workload, from the
        # committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by the copy kernel)
        tk = json.load(open(tf))["kernels"].get(dom_name)
        if tk:
            traffic = tk["traffic_bytes"]
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "bytes_per_launch": dom["bytes"], "mean_ms": round(dom["mean_ms"], 4)}

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as orc
        threads = int(os.environ.get("OMP_NUM_THREADS", str(min(os.cpu_count() or 1, 16))))
        hist_len = args.cpu_iters
        hbase = head[len("mixed_"):] if head.startswith("mixed_") else head
        kwm = dict(ipreflag=2, ilu_method=0, nblocks=256) if hbase == "bj" else dict(ipreflag=1)
        b_h = b_d.numpy()
        if head.startswith("mixed_"):
            # one FP64 correction with its FP32 inner PCG cut to hist_len iterations
            t0 = time.perf_counter()
            orc.pcg_mixed(rp, col, val, b_h, itrmax=1, rtol=1e-30, inner_itrmax=hist_len, inner_rtol=1e-30,
                          **kwm)
            tot = time.perf_counter() - t0
            t0 = time.perf_counter()
            pre = orc.Precon(rp, col, val, kwm["ipreflag"], **{k: v for k, v in kwm.items() if k != "ipreflag"})
            tset = time.perf_counter() - t0
            del pre
            per_it = max(tot - tset, 1e-9) / hist_len
            total_its = H["inner_fp32_iters"]
            cpu = {"value": round(per_it * total_its, 3), "unit": "s", "cores": threads, "kind": "port",
                   "sample": "oracle mixed refinement (%s) on the same 256^3 system: one correction with %d FP32 "
                             "PCG iterations (%.3f s/iteration, preconditioner setup %.1f s excluded), scaled to "
                             "the GPU's %d FP32 iterations" % (head, hist_len, per_it, tset, total_its)}
            no = hist_len
        else:
            t0 = time.perf_counter()
            _, ho, no, _ = orc.pcg(rp, col, val, b_h, itrmax=hist_len, rtol=1e-10, **kwm)
            tset, titer = orc.last_times()
            per_it = titer / no
            cpu = {"value": round(per_it * H["niter"], 3), "unit": "s", "cores": threads, "kind": "port",
                   "sample": "oracle PCG (%s) on the same 256^3 system for %d iterations (%.3f s/iteration, "
                             "setup %.1f s excluded), scaled to the GPU's %d iterations" %
                             (head, no, per_it, tset, H["niter"])}
        log("cpu baseline: %.3f s/it over %d its" % (per_it, no))

    line = {
        "metric": METRIC,
        "value": round(H["time_to_solution_s"], 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(H["time_to_solution_s"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64" if head.startswith("mixed_") else "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D 7-point Poisson %d^3 (n=%d, nnz=%d), x_true~U[-1,1] seed 3, b=A x_true, x0=0, "
                        "rtol 1e-10" % (g, n, nnz),
            "method": {"bj": "FP64 PCG + block-Jacobi ILU(0), 256 row blocks",
                       "pj": "FP64 PCG + point Jacobi",
                       "mixed_bj": "FP64 refinement around FP32 PCG + FP32 block-Jacobi ILU(0), 256 blocks, "
                                   "inner rtol 1e-4",
                       "mixed_pj": "FP64 refinement around FP32 PCG + FP32 point Jacobi, inner rtol 1e-4"}[head],
            "headline": head,
            "speedup_vs_plain_fp64": (round(methods[best_plain]["time_to_solution_s"] / H["time_to_solution_s"], 3)
                                      if best_plain else None),
            "plain_fp64_reference": best_plain,
            "methods": methods,
            "parallelism": "1 GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
