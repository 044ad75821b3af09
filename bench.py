#!/usr/bin/env python3
"""Benchmark: time-to-solution at FP64 relative residual 1e-10 on one MI355X,
learned / mixed-precision PCG against plain FP64 PCG, with HBM GB/s against
the roofline.

Headline workload (C5): 64 test systems of 2D 5-point diffusion on 512^2
cells, conductivity exp(g), g a squared-exponential Gaussian random field
(ell = 16 cells, sigma = 2, field seeds 3048..3111), b ~ N(0,1), x0 = 0.
  learned_apply32 the trained message-passing network predicts every
                  system's BF16 factor L (one batched inference), then the
                  batched FP64 PCG with M = L L^T applied in FP32; the time
                  includes the inference
  learned_refine32 / learned_fp64  FP64 refinement around FP32 PCG / all
                  FP64, same factor
  untrained_apply32, ic0_apply32, ic0_fp64, jacobi_fp64  the same solver
                  with the untrained network / the classical IC(0) factor /
                  point Jacobi (ablation)
  plain_fp64_bj   parcel_dcg's FP64 PCG + block-Jacobi ILU(0) (256 row
                  blocks), one system at a time, solve time only
`value` is learned_apply32's mean time for all 64 systems; the config carries
its speed-up over plain_fp64_bj.
Second workload (C4): 3D 7-point Poisson 256^3 (n = 16,777,216), x_true ~
U[-1,1] seed 3, b = A x_true: FP64 PCG with block-Jacobi ILU(0) / point
Jacobi and the FP64 refinement around FP32 PCG; reported under
config.c4_mixed_vs_plain (or as the headline with --workload c4).
A step is one full solve of the workload with matrices, b and x already in
HBM.  One more solve with HIP events around each kernel gives the per-kernel
means behind `roofline`.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload c5|c4|both]
"""
import argparse
import ctypes
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


def k1_bytes_st7(n, t=8):
    """Fused CG kernel on the symmetric stencil form: the four coefficient
    arrays c, e, n, t (4t per row) and the six vector passes (6t)."""
    return 10 * t * n


def ilu0_apply_bytes(n, nnz_l, nnz_u):
    """Block-Jacobi ILU(0) apply (forward + backward sweep): 12 B per factor
    entry, row ids 4n per sweep, r 8n, s written 8n (forward), s read+write
    16n and u_ii 8n (backward)."""
    return 12 * (nnz_l + nnz_u) + 8 * n + 8 * n + 8 * n + 16 * n + 8 * n


def bj_lines_bytes(n, dilu=False, fused=True, planes=False):
    """Block-Jacobi apply in line-wavefront form (k_bj_lines): forward r 8n
    (and the fused update's q 8n, r written 8n), two couplings 16n (a third,
    the plane neighbour's, 8n when blocks hold several planes), s written 8n
    (D-ILU: the scale 8n); backward s read 8n, the couplings, the scale 8n, s
    written 8n."""
    cpl = 24 * n if planes else 16 * n
    fwd = 8 * n + cpl + 8 * n + (16 * n if fused else 0) + (8 * n if dilu else 0)
    return fwd + 24 * n + cpl


def k1_bytes32(n, nnz):
    """FP32 fused CG kernel: values+columns 8 B/nnz, row pointers 4(n+1),
    x read/write 8n, p_old 4n, s 4n, p_new 4n, q 4n."""
    return 8 * nnz + 4 * (n + 1) + 24 * n


def ilu0_apply_bytes32(n, nnz_l, nnz_u):
    """FP32 block-Jacobi ILU(0) apply: 8 B per factor entry, row ids 4n per
    sweep, r 4n, s written 4n (forward), s read+write 8n and u_ii 4n."""
    return 8 * (nnz_l + nnz_u) + 8 * n + 4 * n + 4 * n + 8 * n + 4 * n


def run_c4(args):
    import numpy as np
    import parcel_hip as ph

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
        if st["spmv_stencil"]:
            # the symmetric seven-point stencil form (4 coefficient arrays)
            kb = k1_bytes_st7(n, 4 if mixed else 8)
            k1name = "cg_spmv_fused_st7" + sfx
        else:
            kb = k1_bytes32(n, nnz) if mixed else k1_bytes(n, nnz)
            k1name = "cg_spmv_fused" + sfx
        kernels = {k1name: {"mean_ms": k1_ms, "calls_per_solve": st["spmv_calls"], "bytes": kb,
                            "GBs": kb / (k1_ms * 1e-3) / 1e9}}
        if base == "bj" and st["bj_lscan"]:
            # the FP32 apply on the grid-plane blocks: line scans of the block
            # IC(0) factor (k_lscan, one workgroup per plane), 40 B per cell
            pb = lscan_bytes(n, 1, 4, 4, 4)
            kernels["bj_ic0_lscan_f32"] = {"mean_ms": pre_ms, "calls_per_solve": st["precon_calls"],
                                           "bytes": pb, "GBs": pb / (pre_ms * 1e-3) / 1e9,
                                           "strips": 256, "lines_per_strip": g}
        elif base == "bj" and st["bj_lines"]:
            # one plane per block: the apply runs as line wavefronts (k_bj_lines)
            pb = bj_lines_bytes(n)
            kernels["bj_ilu0_apply"] = {"mean_ms": pre_ms, "calls_per_solve": st["precon_calls"],
                                        "bytes": pb, "GBs": pb / (pre_ms * 1e-3) / 1e9,
                                        "form": "line wavefronts (k_bj_lines), 88 B per row"}
        elif base == "bj":
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

    # the configured method (FP64 refinement around FP32 PCG + FP32 block-
    # Jacobi ILU(0)) is the line's method whether or not it is the fastest;
    # its plain comparator is FP64 PCG with the same preconditioner
    head = "mixed_bj" if "mixed_bj" in methods else min(methods, key=lambda k: methods[k]["time_to_solution_s"])
    H = methods[head]
    hb = head[len("mixed_"):] if head.startswith("mixed_") else head
    best_plain = hb if hb in methods else None
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
                   "sample": "oracle mixed refinement (%s, its FP32 block ILU(0) applied by sequential "
                             "substitution as the reference's CPU build does) on the same 256^3 system: one "
                             "correction with %d FP32 PCG iterations (%.3f s/iteration, preconditioner setup %.1f s "
                             "excluded), scaled to the GPU's %d FP32 iterations" %
                             (head, hist_len, per_it, tset, total_its)}
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

    return {
        "value": H["time_to_solution_s"],
        "dtype": "f32/f64" if head.startswith("mixed_") else "f64",
        "workload": "3D 7-point Poisson %d^3 (n=%d, nnz=%d), x_true~U[-1,1] seed 3, b=A x_true, x0=0, "
                    "rtol 1e-10" % (g, n, nnz),
        "method": {"bj": "FP64 PCG + block-Jacobi ILU(0), 256 row blocks",
                   "pj": "FP64 PCG + point Jacobi",
                   "mixed_bj": "FP64 refinement around FP32 PCG + FP32 block-Jacobi ILU(0), 256 blocks (the grid "
                               "planes: applied as line scans of the block IC(0) factor, which equals block "
                               "ILU(0) of the symmetric matrix in exact arithmetic), inner rtol 1e-4",
                   "mixed_pj": "FP64 refinement around FP32 PCG + FP32 point Jacobi, inner rtol 1e-4"}[head],
        "headline": head,
        "speedup_vs_plain_fp64": (round(methods[best_plain]["time_to_solution_s"] / H["time_to_solution_s"], 3)
                                  if best_plain else None),
        "plain_fp64_reference": best_plain,
        "fastest_method": min(methods, key=lambda k: methods[k]["time_to_solution_s"]),
        "methods": methods,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def spmv_bytes(n, nnz):
    """FP64 CSR SpMV: values and columns 12 B/nnz, row pointers, x read, y written."""
    return 12 * nnz + 4 * (n + 1) + 16 * n


def gmres_step_bytes(n, nnz, apply_bytes, j):
    """Inner step j (0-based) of right-preconditioned GMRES with MGS: z = M^-1 v_j
    (the block-Jacobi ILU(0) apply, apply_bytes), w = A z, j+1 MGS passes (read
    v_i, read and write w: 24 B each), the norm of w (8 B) and the scaling into
    v_{j+1} (read w, write v: 16 B)."""
    return apply_bytes + spmv_bytes(n, nnz) + 24 * n * (j + 1) + 24 * n


def run_c2(args):
    """3D upwind convection-diffusion 128^3 (cell Peclet 0.5), BJ ILU(0) 64
    blocks, rtol 1e-9: GMRES(30) (MGS) and CA-GMRES(s=8, t=4) with the Newton
    basis from the Leja-ordered Ritz values of the 32 x 32 Hessenberg, TSQR."""
    import numpy as np
    import parcel_hip as ph

    g = args.c2_grid
    n = g ** 3
    t0 = time.perf_counter()
    rp, col, val = ph.convdiff3d(g, g, g, 0.5, (1.0, 0.5, 0.25))
    nnz = int(col.shape[0])
    D = ph.DeviceArray
    xt_d, b_d, x_d = D(ph.uniform(n, 1)), D(n=n), D(n=n)
    log("c2: generated %d^3 n=%d nnz=%d in %.2fs" % (g, n, nnz, time.perf_counter() - t0))
    methods = {}
    variants = {"gmres30": dict(solver=3, GMRES_m=30, GMRES_GSflag=1),
                "cagmres_s8t4_newton": dict(solver=4, CAGMRES_sstep=8, CAGMRES_tstep=4, CAGMRES_basis=2,
                                            CAGMRES_QRflag=3),
                "cagmres_s8t4_newton_reference_step": dict(solver=4, CAGMRES_sstep=8, CAGMRES_tstep=4,
                                                           CAGMRES_basis=1, CAGMRES_QRflag=3)}
    apply_ms = None
    for name, kw in variants.items():
        t = time.perf_counter()
        plan = ph.Plan(rp, col, val, ipreflag=2, ILU_method=0, precon_thblock=64, profile=0, **kw)
        setup = time.perf_counter() - t
        if not methods:
            plan.spmv(xt_d, b_d)
        times = []
        for k in range(args.warmup + args.steps):
            x_d.zero()
            ph.load().parcel_hip_device_synchronize()
            t = time.perf_counter()
            hist, niter, iret = plan.solve(x_d, b_d, 5000, 1e-9)
            if k >= args.warmup:
                times.append(time.perf_counter() - t)
        r_d = D(n=n)
        plan.spmv(x_d, r_d)
        bb = b_d.numpy()
        res = float(np.linalg.norm(bb - r_d.numpy()) / np.linalg.norm(bb))
        st = plan.stats()
        if apply_ms is None:
            apply_ms = plan.time_kernel(2, 20)
            spmv_ms = plan.time_kernel(1, 20)
            # blocks of two grid planes: the line-wavefront apply (k_bj_lines)
            lines_form = bool(st["bj_lines"])
            ab = (bj_lines_bytes(n, fused=False, planes=True) if lines_form
                  else ilu0_apply_bytes(n, st["bj_nnz_l"], st["bj_nnz_u"]))
        methods[name] = {"time_to_solution_s": float(np.mean(times)), "niter": niter, "iret": iret,
                         "final_estimate": float(hist[-1]), "true_rel_res": res, "setup_s": setup,
                         "device_ms": st["solve_device_ms"]}
        log("c2 %s: %.4f s, niter %d, true res %.2e" % (name, methods[name]["time_to_solution_s"], niter, res))
        plan.close()
    # GMRES(30)'s algorithmic bytes over its solve: inner step j of every cycle
    G = methods["gmres30"]
    tot = sum(gmres_step_bytes(n, nnz, ab, i % 30) for i in range(G["niter"]))
    solve_gbs = tot / G["time_to_solution_s"] / 1e9
    ap_gbs = ab / (apply_ms * 1e-3) / 1e9
    kname = ("k_bj_lines (block-Jacobi ILU(0), 64 blocks of two grid planes: forward and backward sweeps as "
             "line wavefronts, one workgroup per block), 88 B per row" if lines_form else
             "k_bj_apply (block-Jacobi ILU(0), 64 blocks: forward and backward level-scheduled sweeps, one "
             "workgroup per block)")
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(ap_gbs, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(ap_gbs / PEAK_HBM_GBS, 4), "traffic": None,
                "bytes_per_launch": ab, "mean_ms": round(apply_ms, 4),
                "solve_level": {"gmres30_algorithmic_bytes": tot, "achieved": round(solve_gbs, 1),
                                "frac": round(solve_gbs / PEAK_HBM_GBS, 4)},
                "fused_cg_spmv_kernel_ms": round(spmv_ms, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as orc
        threads = int(os.environ.get("OMP_NUM_THREADS", str(min(os.cpu_count() or 1, 16))))
        bh = b_d.numpy()
        t = time.perf_counter()
        _, _, no, _ = orc.gmres(rp, col, val, bh, ipreflag=2, ilu_method=0, nblocks=64, itrmax=5000, rtol=1e-9, m=30,
                                gsflag=1)
        tt = time.perf_counter() - t
        tset, titer = orc.last_times() if hasattr(orc, "last_times") else (0.0, tt)
        cpu = {"value": round(tt, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": "oracle GMRES(30) + BJ ILU(0) 64 blocks, the whole C2 solve (%d iterations, "
                         "factorisation included) once" % no}
        log("c2 cpu baseline %.2f s" % tt)
    return {"workload": "3D 7-point first-order upwind convection-diffusion %d^3 (n=%d, nnz=%d), cell Peclet "
                        "max|v| h / (2D) = 0.5, velocity direction (1, 0.5, 0.25), x_true~U[-1,1] seed 1, "
                        "b = A x_true, x0 = 0, rtol 1e-9" % (g, n, nnz),
            "method": "GMRES(30) (MGS) and CA-GMRES(s=8, t=4, Newton basis, TSQR), BJ ILU(0) 64 blocks",
            "value": G["time_to_solution_s"], "methods": methods, "roofline": roofline, "cpu_baseline": cpu}


def ddcg_iter_bytes(n, nnz, nnz_l, nnz_u):
    """One double-double PCG iteration with an FP64 matrix and double-double
    block-Jacobi ILU(0) factors: dd SpMV (12 B/nnz, x and y as hi/lo pairs:
    32n), the dd factor apply (20 B per factor entry: hi, lo, column; row
    data, r in, s out, the backward read-modify-write and the diagonal as
    pairs: 4(n+1)*2 + 96n) and the dd vector passes of the iteration (x, r, p,
    q, s updates and the two dot products: 176n)."""
    return 12 * nnz + 4 * (n + 1) + 32 * n + 20 * (nnz_l + nnz_u) + 8 * (n + 1) + 96 * n + 176 * n


def run_c3(args):
    """3D harmonic-mean diffusion 96^3, log-uniform conductivity in [1e-4, 1e4]
    (seed 2), dd PCG + BJ ILU(0) 32 blocks to rtol 1e-20."""
    import numpy as np
    import parcel_hip as ph

    g = args.c3_grid
    n = g ** 3
    t0 = time.perf_counter()
    rp, col, val = ph.diffusion3d(g, g, g, ph.loguniform(n, 2, 1e-4, 1e4))
    nnz = int(col.shape[0])
    bh, bl = ph.spmv_dd(rp, col, val, ph.uniform(n, 102))
    D = ph.DeviceArray
    bh_d, bl_d, xh_d, xl_d = D(bh), D(bl), D(n=n), D(n=n)
    log("c3: generated %d^3 n=%d nnz=%d in %.2fs" % (g, n, nnz, time.perf_counter() - t0))
    p = ph.default_param(solver=1, ipreflag=2, ILU_method=0, precon_thblock=32)
    lib = ph.load()
    h = ctypes.c_void_p()
    t = time.perf_counter()
    ph._check(lib.parcel_hip_plan_create_dd(ctypes.byref(h), n, n, nnz, 1, ph._ptr(val), None, ph._ptr(rp),
                                            ph._ptr(col), ctypes.byref(p), 1, 2))
    setup = time.perf_counter() - t
    times = []
    hist = np.zeros(20000)
    for k in range(args.warmup + args.steps):
        xh_d.zero()
        xl_d.zero()
        lib.parcel_hip_device_synchronize()
        it, iret = ctypes.c_int(20000), ctypes.c_int(-1)
        t = time.perf_counter()
        ph._check(lib.parcel_hip_plan_solve_dd(h, xh_d.ptr, xl_d.ptr, bh_d.ptr, bl_d.ptr, ctypes.byref(it), 1e-20,
                                               ph._ptr(hist), 2, 2, ctypes.byref(iret)))
        if k >= args.warmup:
            times.append(time.perf_counter() - t)
    st = ph.Stats()
    ph._check(lib.parcel_hip_plan_stats(h, ctypes.byref(st)))
    nl, nu = st.bj_nnz_l, st.bj_nnz_u
    lib.parcel_hip_plan_destroy(h)
    niter = it.value
    tt = float(np.mean(times))
    by = ddcg_iter_bytes(n, nnz, nl, nu) * niter
    gbs = by / tt / 1e9
    log("c3 ddcg: %.4f s, niter %d, final %.3e" % (tt, niter, hist[niter - 1]))
    roofline = {"bound": "hbm", "kernel": "whole dd-PCG iteration (dd SpMV, dd block-Jacobi ILU(0) apply, dd "
                "vector updates and reductions), algorithmic bytes x iterations / solve time",
                "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                "traffic": None, "bytes_per_iteration": ddcg_iter_bytes(n, nnz, nl, nu)}
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as orc
        threads = int(os.environ.get("OMP_NUM_THREADS", str(min(os.cpu_count() or 1, 16))))
        k = args.c3_cpu_iters
        t = time.perf_counter()
        orc.ddcg(rp, col, val, None, bh, bl, ipreflag=2, ilu_method=0, nblocks=32, prec=2, itrmax=k, rtol=1e-30)
        tk = time.perf_counter() - t
        t = time.perf_counter()
        orc.ddcg(rp, col, val, None, bh, bl, ipreflag=2, ilu_method=0, nblocks=32, prec=2, itrmax=1, rtol=1e-30)
        t1 = time.perf_counter() - t
        per_it = (tk - t1) / (k - 1)
        cpu = {"value": round(t1 + per_it * (niter - 1), 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": "oracle dd-PCG + dd BJ ILU(0) 32 blocks on the same system: %d iterations timed (%.4f s per "
                         "iteration after a 1-iteration run with the set-up, %.2f s), scaled to the GPU's %d "
                         "iterations" % (k, per_it, t1, niter)}
        log("c3 cpu baseline %.4f s/it" % per_it)
    return {"workload": "3D 7-point harmonic-mean diffusion %d^3 (n=%d, nnz=%d), kappa log-uniform in [1e-4, 1e4] "
                        "seed 2, x_true~U[-1,1] seed 102, b = A x_true in double-double, x0 = 0, rtol 1e-20 on "
                        "rnorm_hi / bnorm_hi" % (g, n, nnz),
            "method": "double-double PCG, FP64 matrix, double-double BJ ILU(0) 32 blocks",
            "value": tt, "niter": niter, "setup_s": setup, "final_estimate": float(hist[niter - 1]),
            "roofline": roofline, "cpu_baseline": cpu}


def lscan_bytes(n, nsys, t=4, tv=4, tz=None):
    """Factor apply per cell: forward reads r (tv), ls, 1/Ld, a_fwd and writes
    w; backward reads w, 1/Ld, a_bwd, ls and writes z (tz, default tv);
    factor and w in t bytes: 8t + tv + tz (40 B in FP32, 44 B for the FP32
    apply on FP64 vectors with z kept in FP32)."""
    return (8 * t + tv + (tv if tz is None else tz)) * n * nsys


def k1_bytes_batch(n, nnz, nsys, t=8):
    """Batched fused CG kernel: per system values t B/nnz and the vectors of
    k1_bytes (6 vector passes of t bytes); columns and row pointers once per
    launch."""
    return nsys * (t * nnz + 6 * t * n) + 4 * nnz + 4 * (n + 1)


def k1_bytes_stencil(n, nsys, t=8, tz=None):
    """Batched fused CG kernel on the five-point stencil form: per system the
    three coefficient arrays (c, e, n; the w and s couplings are the e and n
    of the neighbouring rows) and the six vector passes of k1_bytes, the
    preconditioned residual z read in tz bytes (default t; 4 when an FP32
    apply keeps it in FP32)."""
    return nsys * (3 * t * n + 5 * t * n + (t if tz is None else tz) * n)


def run_c5(args):
    """Learned-factor mixed PCG on the 64 random-coefficient test systems
    against plain FP64 block-Jacobi ILU(0) PCG on the same GPU."""
    import numpy as np
    import parcel_hip as ph
    import learned as L

    nsys = args.c5_systems
    t0 = time.perf_counter()
    rp, col, vals, b = L.systems(L.TEST_SEED0, nsys)
    n, nnz = b.shape[1], int(col.shape[0])
    log("c5: generated %d systems 512^2 (nnz %d) in %.1fs" % (nsys, nnz, time.perf_counter() - t0))
    D = ph.DeviceArray
    vals_d, b_d = D(vals), D(b)
    x_d = D(np.zeros_like(b))
    y = np.zeros_like(b)
    weights = args.weights
    p = L.load_weights(weights)
    methods = {}

    def true_res(B):
        B.spmv(x_d, y)
        return float(max(np.linalg.norm(b[s] - y[s]) / np.linalg.norm(b[s]) for s in range(nsys)))

    def batched(name, kind, precision, params=None, steps=args.steps, warmup=args.warmup, strips=1,
                ic0_blocks=None):
        t = time.perf_counter()
        B = ph.Batch(rp, col, vals_d, kind, nsys=nsys)
        B.set_inner(args.inner_rtol, 20000)
        if ic0_blocks is not None:
            B.set_ic0_blocks(ic0_blocks)
        elif kind != ph.BATCH_JACOBI:
            B.set_strips(strips)
        setup = time.perf_counter() - t
        net = ph.Mpnn(params) if params is not None else None
        itr = 50 if precision == ph.BATCH_REFINE32 else 20000
        times, niter, iret = [], None, None
        for k in range(warmup + steps):
            x_d.zero()
            ph.load().parcel_hip_device_synchronize()
            t = time.perf_counter()
            if net is not None:
                net.factor(B)
            _, niter, iret = B.solve(x_d, b_d, itr, 1e-10, precision=precision, history=False)
            if k >= warmup:
                times.append(time.perf_counter() - t)
        st = B.stats()
        res = true_res(B)
        B.set_profile(True)
        x_d.zero()
        _, pniter, _ = B.solve(x_d, b_d, itr, 1e-10, precision=precision, history=False)
        pst = B.stats()
        B.close()
        if net is not None:
            net.close()
        m = {"time_to_solution_s": float(np.mean(times)), "steps": len(times), "setup_s": setup, "strips": strips,
             "niter_max": int(niter.max()), "niter_mean": float(niter.mean()), "iret_max": int(iret.max()),
             "true_rel_res_max": res, "device_ms": st["solve_device_ms"],
             "apply_ms": pst["apply_ms"] / max(pst["apply_calls"], 1), "apply_calls": pst["apply_calls"],
             "spmv_ms": pst["spmv_ms"] / max(pst["spmv_calls"], 1), "spmv_calls": pst["spmv_calls"],
             "stencil5": int(pst["stencil5"]), "fallback_systems": int(st["fallback_systems"]),
             # system-iterations of the profiled solve: a converged system's
             # workgroups return at once, so a launch carries only the systems
             # still iterating
             "system_iterations": int(np.sum(pniter))}
        if precision == ph.BATCH_REFINE32:
            m["fp32_batched_iterations"] = st["inner_steps"]
            m["fp32_iterations_per_system_mean"] = st["inner_total"] / nsys
            m["fp64_corrections"] = st["outer_steps"]
        methods[name] = m
        log("c5 %s: %.4f s, its max %d, true res %.2e" % (name, m["time_to_solution_s"], m["niter_max"], res))
        return m

    # headline: FP64 PCG whose preconditioner (the network's BF16 factor) is
    # applied in FP32; then the FP64-refinement form, the all-FP64 form, and
    # the same solver with the untrained network / the classical IC(0) factor
    def ablations(ns):
        batched("learned_apply32_exact_factor", ph.BATCH_FACTOR, ph.BATCH_APPLY32, params=p, steps=1, strips=1)
        batched("learned_refine32", ph.BATCH_FACTOR, ph.BATCH_REFINE32, params=p, steps=1, strips=ns)
        batched("learned_fp64", ph.BATCH_FACTOR, ph.BATCH_FP64, params=p, steps=1, strips=ns)
        batched("untrained_apply32", ph.BATCH_FACTOR, ph.BATCH_APPLY32, params=L.init_params(0), steps=1,
                warmup=0, strips=ns)
        batched("ic0_apply32", ph.BATCH_IC0, ph.BATCH_APPLY32, steps=1, warmup=0, strips=ns)
        batched("ic0_fp64", ph.BATCH_IC0, ph.BATCH_FP64, steps=1, warmup=0, strips=ns)
        batched("jacobi_fp64", ph.BATCH_JACOBI, ph.BATCH_FP64, steps=1, warmup=0)

    # strips per system for the factor solves: the fastest of a one-solve sweep,
    # where counts within 2 % of the fastest (one solve's run-to-run spread)
    # count as a tie, won by the count that keeps more CUs sweeping
    ssweep = {}
    for ns in [int(v) for v in args.c5_strips.split(",")]:
        ssweep[ns] = batched("strips_%d" % ns, ph.BATCH_FACTOR, ph.BATCH_APPLY32, params=p, steps=1,
                             warmup=0, strips=ns)["time_to_solution_s"]
        del methods["strips_%d" % ns]
    fastest = min(ssweep.values())
    ns = max(k for k, v in ssweep.items() if v <= 1.02 * fastest)
    learned = batched("learned_apply32", ph.BATCH_FACTOR, ph.BATCH_APPLY32, params=p, strips=ns)
    learned["strip_sweep_s"] = ssweep
    # the fair plain comparator: plain FP64 PCG with block-Jacobi IC(0) (=
    # block-Jacobi ILU(0) of these symmetric systems in exact arithmetic),
    # the same 64 systems in the same batched launches, timed like the headline
    batched("plain_fp64_bj_batched", ph.BATCH_IC0, ph.BATCH_FP64, ic0_blocks=args.c5_bj_blocks)
    methods["plain_fp64_bj_batched"]["blocks"] = args.c5_bj_blocks
    if args.c5_ablations:
        ablations(ns)
    return finish_c5(args, rp, col, vals, b, n, nnz, nsys, p, methods, learned, weights)


def finish_c5(args, rp, col, vals, b, n, nnz, nsys, p, methods, learned, weights):
    import numpy as np
    import parcel_hip as ph
    D = ph.DeviceArray

    # plain FP64 block-Jacobi ILU(0) PCG, parcel_dcg's solver, one system at a time;
    # the block count is the fastest of a sweep on the first 4 systems
    sweep = {}
    for nbc in [int(v) for v in args.c5_blocks.split(",")]:
        tt = 0.0
        for s in range(min(4, nsys)):
            plan = ph.Plan(rp, col, vals[s], ipreflag=2, ILU_method=0, precon_thblock=nbc, profile=0)
            xs, bs = D(n=n), D(b[s])
            ph.load().parcel_hip_device_synchronize()
            t = time.perf_counter()
            _, niter, _ = plan.solve(xs, bs, 20000, 1e-10)
            tt += time.perf_counter() - t
            plan.close()
        sweep[nbc] = tt / min(4, nsys)
        log("c5 plain BJ(%d): %.3f s per system (4-system sweep)" % (nbc, sweep[nbc]))
    nb = min(sweep, key=sweep.get)
    solve_t, setup_t, its, res, tres = [], [], [], [], []
    nplain = min(nsys, args.c5_plain_systems) if args.c5_plain_systems > 0 else nsys
    for s in range(nplain):
        t = time.perf_counter()
        plan = ph.Plan(rp, col, vals[s], ipreflag=2, ILU_method=0, precon_thblock=nb, profile=0)
        t1 = time.perf_counter()
        xs, bs = D(n=n), D(b[s])
        ph.load().parcel_hip_device_synchronize()
        t2 = time.perf_counter()
        hist, niter, iret = plan.solve(xs, bs, 20000, 1e-10)
        t3 = time.perf_counter()
        solve_t.append(t3 - t2)
        setup_t.append(t1 - t)
        its.append(niter)
        if s < 4:
            r_d = D(n=n)
            plan.spmv(xs, r_d)
            tres.append(float(np.linalg.norm(b[s] - r_d.numpy()) / np.linalg.norm(b[s])))
        res.append(float(hist[-1]))
        plan.close()
    # (with --c5-plain-systems below nsys, a profiling run: the total is scaled)
    plain = {"time_to_solution_s": float(np.sum(solve_t)) * nsys / nplain, "systems_timed": nplain,
             "setup_s_total": float(np.sum(setup_t)) * nsys / nplain,
             "block_sweep_s_per_system": sweep,
             "niter_max": int(max(its)), "niter_mean": float(np.mean(its)), "final_recurrence_res_max": max(res),
             "true_rel_res_max_first4": max(tres), "blocks": nb}
    methods["plain_fp64_bj"] = plain
    log("c5 plain FP64 BJ(%d): %.3f s solve (setup %.2f s), its mean %.0f" % (nb, plain["time_to_solution_s"],
                                                                           plain["setup_s_total"],
                                                                           plain["niter_mean"]))
    # dominant kernel of the learned solve (the larger total time of the two
    # big iteration kernels; the other one is reported beside it)
    # algorithmic bytes per launch = per-system bytes x the system-iterations
    # of the profiled solve / launches (converged systems' workgroups return
    # at once and the last chunk of launches finds every system done)
    sysit = learned["system_iterations"]
    lscan_k = ("k_lscan<float,double,2,4,FULL,float> (FP32 factor apply on FP64 vectors, z kept in FP32, 4 "
               "wavefronts per line, %d strip(s) per system)" % learned["strips"], learned["apply_ms"],
               lscan_bytes(n, 1, 4, 8, 4) * sysit / max(learned["apply_calls"], 1), learned["apply_calls"])
    if learned.get("stencil5"):
        spmv_k = ("k_cg_spmv<double,4,St5,float> (batched, five-point stencil, z in FP32)", learned["spmv_ms"],
                  k1_bytes_stencil(n, 1, 8, 4) * sysit / max(learned["spmv_calls"], 1), learned["spmv_calls"])
    else:
        spmv_k = ("k_cg_spmv<double> (batched)", learned["spmv_ms"],
                  (k1_bytes_batch(n, nnz, 1, 8) - 4 * nnz - 4 * (n + 1)) * sysit / max(learned["spmv_calls"], 1)
                  + 4 * nnz + 4 * (n + 1), learned["spmv_calls"])
    ap_total = learned["apply_ms"] * learned["apply_calls"]
    sp_total = learned["spmv_ms"] * learned["spmv_calls"]
    dom, other = (lscan_k, spmv_k) if ap_total >= sp_total else (spmv_k, lscan_k)
    tf = os.path.join(ROOT, "profiles", "batched_pcg_c5_grf512_traffic.json")
    tks = {}
    if nsys == 64 and n == 512 * 512 and os.path.exists(tf):
        # HBM bytes per launch of the same kernel on the same workload, from the
        # committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by the copy kernel)
        tks = json.load(open(tf))["kernels"]

    def line(k):
        kname, ms, by, calls = k
        ach = by / (ms * 1e-3) / 1e9
        tk = tks.get(kname)
        return {"bound": "hbm", "kernel": kname, "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": tk["traffic_bytes"] if tk else None,
                "bytes_per_launch": int(by), "mean_ms": round(ms, 4), "launches": calls,
                "active_systems_per_launch": round(sysit / max(calls, 1), 2)}

    roofline = line(dom)
    roofline["next_kernel"] = line(other)
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as orc
        threads = int(os.environ.get("OMP_NUM_THREADS", str(min(os.cpu_count() or 1, 16))))
        ks = min(args.cpu_systems, nsys)
        runs = []
        for rep in range(1 + args.cpu_reps):  # one warm-up run, then the timed ones
            t = time.perf_counter()
            cits = []
            for s in range(ks):
                _, _, no, io = orc.pcg(rp, col, vals[s], b[s], ipreflag=2, ilu_method=0, nblocks=args.c5_bj_blocks,
                                       itrmax=20000, rtol=1e-10)
                cits.append(no)
            if rep:
                runs.append(time.perf_counter() - t)
        med = float(np.median(runs))
        cpu = {"value": round(med * nsys / ks, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": "oracle FP64 PCG + BJ ILU(0) (%d blocks; set-up included) solving test systems 0..%d to "
                         "1e-10 (%s iterations): median of %d runs %.3f s (runs %s), times %d/%d for the 64-system "
                         "workload" % (args.c5_bj_blocks, ks - 1, cits, len(runs), med,
                                       [round(v, 3) for v in runs], nsys, ks)}
    meta = {}
    mp = os.path.splitext(weights)[0] + ".json"
    if os.path.exists(mp):
        meta = {k: v for k, v in json.load(open(mp)).items() if k != "loss_curve_every50"}
    return {
        "value": learned["time_to_solution_s"],
        "dtype": "bf16/f32/f64",
        "workload": "2D 5-point diffusion 512^2, kappa=exp(GRF, SE ell=16 sigma=2), %d test systems (field seeds "
                    "3048..%d), b~N(0,1), x0=0, FP64 rtol 1e-10 per system" % (nsys, 3048 + nsys - 1),
        "method": "learned factor (6x32 MPNN, BF16 MFMA inference, BF16 L) applied in FP32 inside batched FP64 "
                  "PCG (64 systems per launch), factor solves on %d strip(s) of grid lines per system "
                  "(M = diag(L_k L_k^T)); time includes the network inference" % learned["strips"],
        "headline": "learned_apply32",
        "speedup_vs_plain_fp64": round(methods["plain_fp64_bj_batched"]["time_to_solution_s"] /
                                       learned["time_to_solution_s"], 3),
        "plain_fp64_reference": "plain_fp64_bj_batched: FP64 PCG + block-Jacobi IC(0) with %d blocks of 16 lines "
                                "(block-Jacobi ILU(0) of these symmetric systems in exact arithmetic; its "
                                "iteration counts equal parcel_dcg's), all 64 systems in the same batched "
                                "launches as the headline" % args.c5_bj_blocks,
        "speedup_vs_ic0_fp64": (round(methods["ic0_fp64"]["time_to_solution_s"] / learned["time_to_solution_s"], 3)
                                if "ic0_fp64" in methods else None),
        "speedup_vs_parcel_dcg_one_system_at_a_time": round(plain["time_to_solution_s"] /
                                                            learned["time_to_solution_s"], 3),
        "parcel_dcg_reference": "plain_fp64_bj: parcel_dcg FP64 PCG + BJ ILU(0) %d blocks (fastest of the sweep), "
                                "one system at a time as the reference API solves them, solve time only "
                                "(factorisation excluded)" % nb,
        "methods": methods,
        "training": meta,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c5", "c4", "c2", "c3", "both", "all"], default="all",
                    help="both: C5 and C4; all: C5, C4, C2 and C3")
    ap.add_argument("--c2-grid", type=int, default=128)
    ap.add_argument("--c3-grid", type=int, default=96)
    ap.add_argument("--c3-cpu-iters", type=int, default=60)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--methods", default="bj,pj,mixed_bj,mixed_pj")
    ap.add_argument("--c5-systems", type=int, default=64)
    ap.add_argument("--c5-blocks", default="8,16,32,64,256",
                    help="block counts of the plain FP64 BJ ILU(0) baseline sweep (fastest is used)")
    ap.add_argument("--inner-rtol", type=float, default=1e-4)
    ap.add_argument("--c5-strips", default="1,2,4,8", help="strips per system of the factor solves (sweep)")
    ap.add_argument("--c5-ablations", type=int, default=1, help="0: only the headline and the plain baseline")
    ap.add_argument("--c5-plain-systems", type=int, default=0,
                    help="time the plain baseline on this many systems and scale (profiling runs; 0: all)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "parcel-legacy_amd", "weights", "mpnn_grf512.npy"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--cpu-systems", type=int, default=4, help="C5 CPU baseline: systems solved to convergence")
    ap.add_argument("--cpu-reps", type=int, default=5, help="C5 CPU baseline: timed runs (median), after a warm-up")
    ap.add_argument("--c5-bj-blocks", type=int, default=32,
                    help="blocks of the batched block-Jacobi comparator (a divisor of 512)")
    ap.add_argument("--calib", action="store_true",
                    help="also run one plain 2 GiB device-to-device copy (reference for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one GPU holds one rank's whole subdomain); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import parcel_hip as ph

    ph.init(0)
    c5 = c4 = c2 = c3 = None
    if args.workload in ("c5", "both", "all"):
        if os.path.exists(args.weights):
            c5 = run_c5(args)
        else:
            log("c5 skipped: no trained weights at %s" % args.weights)
    if args.workload in ("c4", "both", "all"):
        c4 = run_c4(args)
    if args.workload in ("c2", "all"):
        c2 = run_c2(args)
    if args.workload in ("c3", "all"):
        c3 = run_c3(args)
    H = c5 if c5 is not None else c4
    if H is None:
        H = {"workload": (c2 or c3)["workload"], "method": (c2 or c3)["method"], "headline": None,
             "speedup_vs_plain_fp64": None, "plain_fp64_reference": None, "methods": (c2 or {}).get("methods"),
             "value": (c2 or c3)["value"], "dtype": "f64" if c2 else "dd", "roofline": (c2 or c3)["roofline"],
             "cpu_baseline": (c2 or c3)["cpu_baseline"]}
    cfg = {"workload": H["workload"], "method": H["method"], "headline": H["headline"],
           "speedup_vs_plain_fp64": H["speedup_vs_plain_fp64"], "plain_fp64_reference": H["plain_fp64_reference"],
           "methods": H["methods"], "parallelism": "1 GPU"}
    if c5 is not None:
        cfg["training"] = c5["training"]
        for k in ("speedup_vs_ic0_fp64", "speedup_vs_parcel_dcg_one_system_at_a_time", "parcel_dcg_reference"):
            cfg[k] = c5[k]
        if c4 is not None:
            cfg["c4_mixed_vs_plain"] = {k: c4[k] for k in ("workload", "method", "headline", "value",
                                                            "speedup_vs_plain_fp64", "plain_fp64_reference",
                                                            "fastest_method", "methods", "roofline",
                                                            "cpu_baseline")}
    if c2 is not None and H is not c2:
        cfg["c2_convdiff_gmres"] = c2
    if c3 is not None:
        cfg["c3_ddcg"] = c3
    line = {
        "metric": METRIC,
        "value": round(H["value"], 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(H["value"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": H["dtype"],
        "data": "synthetic",
        "config": cfg,
        "roofline": H["roofline"],
        "cpu_baseline": H["cpu_baseline"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
