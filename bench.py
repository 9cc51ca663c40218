"""Headline benchmark: gz + FTG forward and adjoint on one MI355X.

Workload: 128x128x64 cubic cells of 20 m (1,048,576 cells), 64x64 sites at
z = -30 m on cell centres, density 2670 + N(0, 20) + five dipping dikes
(+/-300 kg/m^3), two background layers.  One step = one forward (gz + the 9 FTG
values per site) plus one matrix-free adjoint (gradient of a least-squares misfit
over all 10 values per site) with all inputs resident in HBM.

value = site-cell pairs per second of a forward+adjoint step (4096 x 1048576 pairs
per step) for the mixed-precision path; the plain FP64 path is timed beside it.
The CPU baseline is the oracle (the OpenMP restatement of the reference's OMP imps:
scalar and tensor calculators run separately, as jif3D does) on a bounded sample of
sites, scaled linearly (work per site is identical).

The same line carries "storedj": the batched stored-J products (J M and J^T R with 64
right-hand sides on a 7168 x 131072 FP32 J) in FP32, BF16 and BF16-refined MFMA modes,
with HBM and matrix-core fractions.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64|both]
                    [--workload all|headline|storedj|configs]

`--workload configs` prints instead one table row per benchmark configuration (gz
forward; FTG forward; gz+FTG adjoint with a finite-difference check; magnetic total-field
forward + adjoint; magnetic vertical gradient forward + adjoint), each with the oracle's CPU
rate beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("site·cell box-term evals/s fwd+adjoint (mixed vs FP64) at stated error; "
          "% vector peak")
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12}  # MI355X vector FP64 / FP32 (flop/s), published
# FP32 vector rate without packed-FP32 instructions, which this build disables: 256 CUs x
# 4 SIMDs x 16 lanes x 2 flop (FMA) x 2.4 GHz; the published 157.3 TFLOP/s counts packed FP32
FP32_UNPACKED = 78.6e12


def fe_per_pair(rho_c):
    """Flop-equivalent contract per site-cell pair for gz + FTG, forward only."""
    return 133.0 * rho_c + 70.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# library timer names -> kernel symbols in the rocprofv3 summaries (cubic mesh)
KERNEL_SYMBOL = {
    "forward_corners_fp64_gz_ftg": "k_forward_corners<7, 1>",
    "adjoint_corners_fp64_gz_ftg": "k_adjoint_corners<7, 1>",
    "far_forward_fp32_gz_ftg": "k_far_forward<7, 0, 1>",
    "far_adjoint_fp32_gz_ftg": "k_far_adjoint<7, 0, 1>",
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gz_ftg_128x128x64_hbm_traffic.md")
VALU_FILE = os.path.join(ROOT, "profiles", "gz_ftg_128x128x64_valu_counters.md")


def profiled_traffic(kname):
    """HBM bytes per dispatch of a kernel from the committed counter summary, or None."""
    sym = KERNEL_SYMBOL.get(kname)
    if sym is None or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0] == f"`{sym}`":
            return float(cells[3]) * 1e6
    return None


def profiled_flops_per_pair(kname):
    """FP32 + FP64 FLOPs per site-cell pair a kernel executes, from the committed counter
    summary (SQ_INSTS_VALU_FLOPS_* per dispatch x 64 lanes / pairs per dispatch), or None."""
    sym = KERNEL_SYMBOL.get(kname)
    if sym is None or not os.path.exists(VALU_FILE):
        return None
    for line in open(VALU_FILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == f"`{sym}`":
            return float(cells[5])
    return None


def run_gpu(args, model, data, prec_name):
    from jif3d_amd import _native as nat
    prec = nat.MIXED if prec_name == "mixed" else nat.FP64
    kind = nat.KIND_SCALAR_TENSOR
    ctx = nat.Context(0)
    ctx.configure(model, data)
    S, N = data.nsites, model.ncells
    mv = ctx.to_device(model.GetModelParameters())
    out = ctx.alloc(S * 10)
    rng = np.random.default_rng(3)
    resid = ctx.to_device(rng.normal(size=S * 10) * 1e-9)
    grad = ctx.alloc(ctx.n_params())

    def step():
        ctx.forward(kind, prec, mv, out)
        ctx.adjoint(kind, prec, mv, resid, grad)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    log(f"[bench] {prec_name}: warmup done")
    ctx.profile_enable(True)
    ctx.profile_reset()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile()
    ctx.profile_enable(False)
    ms = (t1 - t0) * 1e3 / args.steps
    log(f"[bench] {prec_name}: {ms:.2f} ms/step")
    result = {"ms_per_step": ms, "pairs_per_s": S * N / (ms * 1e-3), "kernels": prof,
              "out": out.to_host(), "grad": grad.to_host()}
    return result


MFMA_PEAK = {"fp32": 157.3e12, "bf16": 2.5e15, "fp64": 78.6e12}  # MI355X dense matrix peaks
HBM_PEAK = 8.0e12
PCIE_PEAK = 64.0e9  # PCIe 5.0 x16, one direction


def run_storedj(args):
    """Batched stored-J: 64x64x32 cells x 32x32 sites x 7 components (scalar+FTG), J in
    HBM as FP32 (3.76 GB); 64 trial models (J M) and 64 residual vectors (J^T R) per
    product, each precision timed over `steps` products after `warmup`."""
    from jif3d_amd import _native as nat, synthetic
    m, d, M, Rr = synthetic.config3()
    m.SetBackgroundThicknesses([])
    m.SetBackgroundDensities([])
    ctx = nat.Context(0)
    ctx.configure(m, d)
    ctx.profile_enable(True)
    ctx.profile_reset()
    nrhs = M.shape[0]
    Md = ctx.to_device(M.astype(np.float32))
    Rd = ctx.to_device(Rr.astype(np.float32))
    # the J kept in page-locked host memory and streamed over PCIe in blocks of 16384 cells
    # (J larger than HBM), FP32 products, first in the process: after the multi-GB HBM-resident
    # J below had been built and released in the same process its first passes ran at about
    # half the PCIe rate; one warm-up pair, the median of three timed pairs
    t0 = time.perf_counter()
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64, host_block_cells=16384)
    ctx.synchronize()
    host = {"build_ms": (time.perf_counter() - t0) * 1e3, "block_cells": 16384}
    R, N = ctx.storedj_rows(), ctx.storedj_cols()
    Y = ctx.alloc((nrhs, R), np.float32)
    G = ctx.alloc((nrhs, N), np.float32)
    jbytes = 4.0 * R * N
    flops = 2.0 * R * N * nrhs
    host_t = {"J_M": [], "JT_R": []}
    for step in range(4):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.storedj_apply(nat.GEMM_FP32, Md, Y)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.storedj_apply_t(nat.GEMM_FP32, Rd, G)
        ctx.synchronize()
        t2 = time.perf_counter()
        if step > 0:
            host_t["J_M"].append(t1 - t0)
            host_t["JT_R"].append(t2 - t1)
    for key, ts in host_t.items():
        dt = float(np.median(ts))
        host[key] = {"ms": dt * 1e3, "J_stream_GB_s": jbytes / dt / 1e9,
                     "pcie_frac": jbytes / dt / PCIE_PEAK}
    host_out = (Y.to_host().astype(np.float64), G.to_host().astype(np.float64))
    log(f"[bench] stored J in host memory: {host['J_M']['ms']:.1f} ms J M, "
        f"{host['JT_R']['ms']:.1f} ms J^T R")
    ctx.storedj_release()
    ctx.profile_reset()
    t0 = time.perf_counter()
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64)
    ctx.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    res = {"workload": f"J {R} x {N} FP32 ({jbytes / 1e9:.2f} GB), {nrhs} right-hand sides",
           "build_ms": build_ms}
    ref = None
    for name, prec in (("fp32", nat.GEMM_FP32), ("bf16", nat.GEMM_BF16),
                       ("bf16_refined", nat.GEMM_BF16_REFINED)):
        for _ in range(args.warmup):
            ctx.storedj_apply(prec, Md, Y)
            ctx.storedj_apply_t(prec, Rd, G)
        ctx.synchronize()
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.storedj_apply(prec, Md, Y)
            ctx.storedj_apply_t(prec, Rd, G)
        ctx.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / args.steps
        prof = ctx.profile()
        peak = MFMA_PEAK["fp32" if prec == nat.GEMM_FP32 else "bf16"]
        entry = {"ms_per_pair_of_products": wall}
        for key, kname in (("J_M", f"storedj_jm_{name}"), ("JT_R", f"storedj_jtr_{name}")):
            # the BF16 J M that read the packed BF16 copy of J (2 bytes per element) is
            # timed under its own name; every other product streams the FP32 J
            packed = kname + "_packed"
            if packed in prof:
                kname = packed
            kms, n = prof[kname]
            avg = kms / n * 1e-3
            jb = jbytes / 2 if kname.endswith("_packed") else jbytes
            entry[key] = {"kernel": kname, "kernel_ms": avg * 1e3,
                          "J_bytes_per_element": jb / (R * N),
                          "J_stream_GB_s": jb / avg / 1e9, "hbm_frac": jb / avg / HBM_PEAK,
                          "TFLOP_s": flops / avg / 1e12, "mfma_frac": flops / avg / peak}
        y, g = Y.to_host().astype(np.float64), G.to_host().astype(np.float64)
        if ref is None:
            ref = (y, g)
        else:
            entry["rel_err_vs_fp32"] = {
                "J_M": float(np.max(np.abs(y - ref[0])) / np.max(np.abs(ref[0]))),
                "JT_R": float(np.max(np.abs(g - ref[1])) / np.max(np.abs(ref[1])))}
        res[name] = entry
        log(f"[bench] stored J {name}: {wall:.3f} ms per J M + J^T R")
    ctx.storedj_release()
    # the same J in FP64 (the reference's precision) on the FP64 matrix cores
    t0 = time.perf_counter()
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64, fp64=True)
    ctx.synchronize()
    f64 = {"build_ms": (time.perf_counter() - t0) * 1e3}
    Md64, Rd64 = ctx.to_device(M.astype(np.float64)), ctx.to_device(Rr.astype(np.float64))
    Y64, G64 = ctx.alloc((nrhs, R)), ctx.alloc((nrhs, N))
    for _ in range(args.warmup):
        ctx.storedj_apply_f64(Md64, Y64)
        ctx.storedj_apply_t_f64(Rd64, G64)
    ctx.synchronize()
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.storedj_apply_f64(Md64, Y64)
        ctx.storedj_apply_t_f64(Rd64, G64)
    ctx.synchronize()
    f64["ms_per_pair_of_products"] = (time.perf_counter() - t0) * 1e3 / args.steps
    prof = ctx.profile()
    jb64 = 8.0 * R * N
    for key, kname in (("J_M", "storedj_jm_fp64"), ("JT_R", "storedj_jtr_fp64")):
        kms, n = prof[kname]
        avg = kms / n * 1e-3
        f64[key] = {"kernel": kname, "kernel_ms": avg * 1e3, "J_bytes_per_element": 8.0,
                    "J_stream_GB_s": jb64 / avg / 1e9, "hbm_frac": jb64 / avg / HBM_PEAK,
                    "TFLOP_s": flops / avg / 1e12, "mfma_frac": flops / avg / MFMA_PEAK["fp64"]}
    y64, g64 = Y64.to_host(), G64.to_host()
    f64["rel_diff_vs_fp32"] = {
        "J_M": float(np.max(np.abs(y64 - ref[0])) / np.max(np.abs(y64))),
        "JT_R": float(np.max(np.abs(g64 - ref[1])) / np.max(np.abs(g64)))}
    res["fp64"] = f64
    log(f"[bench] stored J fp64: {f64['ms_per_pair_of_products']:.3f} ms per J M + J^T R")
    for a in (Md64, Rd64, Y64, G64):
        a.free()
    ctx.storedj_release()
    y, g = host_out
    host["rel_diff_vs_hbm_resident"] = {
        "J_M": float(np.max(np.abs(y - ref[0])) / np.max(np.abs(ref[0]))),
        "JT_R": float(np.max(np.abs(g - ref[1])) / np.max(np.abs(ref[1])))}
    res["host_resident_fp32"] = host
    return res


FE_FWD = {"gz": (55.0, 10.0), "ftg": (106.0, 60.0), "gz+ftg": (133.0, 70.0),
          "magnetic": (106.0, 72.0),
          "maggrad": (212.0, 122.0)}  # flop-equivalents per pair: a * rho_c + b (maggrad:
#                                       two tensor box terms, the site and 1 m below)


def run_configs(args):
    """The survey's five benchmark configurations, one table row each: GPU pairs/s and
    flop-equivalent rate (FP64 and mixed where both exist), error against the FP64 path
    and against the oracle's long double, and the oracle's CPU rate on a sample of sites
    in the same process.  Forward-only rows count forward pairs; adjoint rows adjoint
    pairs; the fwd+adj rows count a pair once per forward + adjoint."""
    from jif3d_amd import _native as nat, synthetic
    from oracle import oracle as orc
    reps = max(3, args.steps)
    rows = []

    def timed(ctx, fn):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) / reps

    def cpu_rate(fn, nsites_total, budget=2.0):
        t0 = time.perf_counter()
        fn(1)
        t1 = time.perf_counter() - t0
        n = int(max(1, min(nsites_total, budget / max(t1, 1e-4))))
        t0 = time.perf_counter()
        fn(n)
        return n, (time.perf_counter() - t0)

    def row(cfg, what, model, data, times, fe_key, errors, cpu, passes=1):
        S, N = data.nsites, model.ncells
        rho_c = (model.nx + 1) * (model.ny + 1) * (model.nz + 1) / N
        a, b = FE_FWD[fe_key]
        fe = passes * (a * rho_c + b)  # forward and adjoint count the same
        r = {"config": cfg, "what": what, "cells": N, "sites": S, "rho_c": rho_c,
             "fe_per_pair": fe}
        for prec, t in times.items():
            pairs = S * N / t
            peak = PEAK["fp64"] if prec == "fp64" else PEAK["fp32"]
            r[prec] = {"ms": t * 1e3, "pairs_per_s": pairs, "fe_per_s": pairs * fe,
                       "frac_of_vector_peak": pairs * fe / peak}
        if "fp64" in times and "mixed" in times:
            r["speedup_mixed_vs_fp64"] = times["fp64"] / times["mixed"]
        r["errors"] = errors
        n, t = cpu
        r["cpu"] = {"pairs_per_s": n * N / t, "cores": orc.num_threads(),
                    "sample_sites": n}
        rows.append(r)
        log(f"[bench] config {cfg} {what}: " + ", ".join(
            f"{p} {v['ms']:.3f} ms" for p, v in r.items() if isinstance(v, dict) and "ms" in v))

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    def comp_rel(a, b, k):
        a, b = a.reshape(-1, k), b.reshape(-1, k)
        floor = 1e-6 * np.max(np.abs(b))
        return float(max(np.max(np.abs(a[:, c] - b[:, c])) / max(np.max(np.abs(b[:, c])), floor)
                         for c in range(k)))

    # config 0: gz forward, FP64, against the oracle on every site
    m, d = synthetic.config0()
    ctx = nat.Context(0)
    ctx.configure(m, d)
    mv, x = m.GetModelParameters(), d.positions
    dm, out = ctx.to_device(mv), ctx.alloc(d.nsites)
    t64 = timed(ctx, lambda: ctx.forward(nat.KIND_SCALAR, nat.FP64, dm, out))
    g = out.to_host()
    errs = {"fp64_vs_oracle_fp64": rel(g, orc.calculate(orc.SCALAR, m, mv, x)),
            "fp64_vs_long_double": rel(g, orc.calculate_ld(orc.SCALAR, m, mv, x))}
    cpu = cpu_rate(lambda n: orc.calculate(orc.SCALAR, m, mv, tuple(c[:n] for c in x)),
                   d.nsites)
    row(0, "gz forward", m, d, {"fp64": t64}, "gz", errs, cpu)

    # config 1: FTG forward, FP64 and mixed
    m, d = synthetic.config1()
    ctx = nat.Context(0)
    ctx.configure(m, d)
    mv, x = m.GetModelParameters(), d.positions
    dm, out = ctx.to_device(mv), ctx.alloc(d.nsites * 9)
    res, times = {}, {}
    for name, prec in (("fp64", nat.FP64), ("mixed", nat.MIXED)):
        times[name] = timed(ctx, lambda: ctx.forward(nat.KIND_TENSOR, prec, dm, out))
        res[name] = out.to_host()
    sub = np.random.default_rng(11).choice(d.nsites, 24, replace=False)
    xs = tuple(c[sub] for c in x)
    ld = orc.calculate_ld(orc.TENSOR, m, mv, xs)
    errs = {"mixed_vs_fp64_per_component": comp_rel(res["mixed"], res["fp64"], 9),
            "fp64_vs_long_double_24_sites": comp_rel(res["fp64"].reshape(-1, 9)[sub], ld, 9),
            "mixed_vs_long_double_24_sites": comp_rel(res["mixed"].reshape(-1, 9)[sub], ld, 9)}
    cpu = cpu_rate(lambda n: orc.calculate(orc.TENSOR, m, mv, tuple(c[:n] for c in x)),
                   d.nsites)
    row(1, "FTG forward", m, d, times, "ftg", errs, cpu)

    # config 2: gz + FTG adjoint (gradient of sum |d - d_obs|^2)
    m, d = synthetic.config2()
    ctx = nat.Context(0)
    ctx.configure(m, d)
    kind = nat.KIND_SCALAR_TENSOR
    true = m.GetModelParameters()
    start = true.copy()
    start[:m.ncells] = 2670.0
    rng = np.random.default_rng(21)
    d_true = ctx.forward(kind, nat.FP64, true)
    scale = np.abs(d_true.reshape(-1, 10)).max(axis=0)
    d_obs = d_true + (0.01 * rng.normal(size=d_true.size).reshape(-1, 10) * scale).reshape(-1)
    r = ctx.forward(kind, nat.FP64, start) - d_obs
    dm, dr, grad = ctx.to_device(start), ctx.to_device(r), ctx.alloc(ctx.n_params())
    res, times = {}, {}
    for name, prec in (("fp64", nat.FP64), ("mixed", nat.MIXED)):
        times[name] = timed(ctx, lambda: ctx.adjoint(kind, prec, dm, dr, grad))
        res[name] = grad.to_host()
    cells = np.random.default_rng(22).choice(m.ncells, 20, replace=False)
    fd = np.empty(cells.size)
    for i, c in enumerate(cells):
        p_, q_ = start.copy(), start.copy()
        p_[c] += 0.1
        q_[c] -= 0.1
        dd = ctx.forward(kind, nat.FP64, p_) - ctx.forward(kind, nat.FP64, q_)
        fd[i] = 2.0 * np.dot(r, dd) / 0.2
    gmax = np.max(np.abs(res["fp64"]))
    errs = {"mixed_vs_fp64": rel(res["mixed"], res["fp64"]),
            "fp64_vs_central_fd_20_cells": float(np.max(np.abs(res["fp64"][cells] - fd)) / gmax),
            "mixed_vs_central_fd_20_cells": float(np.max(np.abs(res["mixed"][cells] - fd)) / gmax)}
    x = d.positions
    rr = r.reshape(-1, 10)

    def cpu_adj(n):
        xs = tuple(c[:n] for c in x)
        orc.lqderivative(orc.SCALAR, m, start, xs, rr[:n, 0].copy())
        orc.lqderivative(orc.TENSOR, m, start, xs, rr[:n, 1:].reshape(-1).copy())
    cpu = cpu_rate(cpu_adj, d.nsites)
    row(2, "gz+FTG adjoint", m, d, times, "gz+ftg", errs, cpu)

    # config 4: magnetic total-field anomaly forward + adjoint
    m, d, field = synthetic.config4_magnetic()
    ctx = nat.Context(0)
    ctx.configure(m, d, field)
    mv, x = m.GetModelParameters(), d.positions
    kind = nat.KIND_MAGNETIC
    sub = np.random.default_rng(31).choice(d.nsites, 6, replace=False)
    xs = tuple(c[sub] for c in x)
    b64 = ctx.forward(kind, nat.FP64, mv).reshape(-1, 3)
    ld = orc.calculate_ld(orc.MAGNETIC, m, mv, xs, field=field)
    ctx.set_transform(nat.TF_TOTAL_FIELD_ANOMALY, field[0], field[1])
    dm, out = ctx.to_device(mv), ctx.alloc(d.nsites)
    dr, grad = ctx.to_device(np.random.default_rng(32).normal(size=d.nsites)), ctx.alloc(ctx.n_params())
    res, times = {}, {}
    for name, prec in (("fp64", nat.FP64), ("mixed", nat.MIXED)):
        def step(prec=prec):
            ctx.forward(kind, prec, dm, out)
            ctx.adjoint(kind, prec, dm, dr, grad)
        times[name] = timed(ctx, step)
        res[name] = (out.to_host(), grad.to_host())
    errs = {"mixed_vs_fp64_tfa": rel(res["mixed"][0], res["fp64"][0]),
            "mixed_vs_fp64_gradient": rel(res["mixed"][1], res["fp64"][1]),
            "fp64_vs_long_double_B_6_sites": comp_rel(b64[sub], ld, 3)}
    rtf = np.random.default_rng(33).normal(size=d.nsites)

    def cpu_mag(n):
        xn = tuple(c[:n] for c in x)
        orc.calculate(orc.MAGNETIC, m, mv, xn, field=field,
                      transform=orc.TF_TOTAL_FIELD_ANOMALY, tf_angles=field[:2])
        orc.lqderivative(orc.MAGNETIC, m, mv, xn, rtf[:n], field=field,
                         transform=orc.TF_TOTAL_FIELD_ANOMALY, tf_angles=field[:2])
    cpu = cpu_rate(cpu_mag, d.nsites, budget=4.0)
    row(4, "magnetic TFA forward + adjoint", m, d, times, "magnetic", errs, cpu, passes=2)

    # config 4 geometry, vertical magnetic gradient (two-point FP64; slab form in mixed)
    ctx.set_transform(nat.TF_NONE, 0.0, 0.0)
    kind = nat.KIND_MAGNETIC_GRADIENT
    res, times = {}, {}
    for name, prec in (("fp64", nat.FP64), ("mixed", nat.MIXED)):
        def step(prec=prec):
            ctx.forward(kind, prec, dm, out)
            ctx.adjoint(kind, prec, dm, dr, grad)
        times[name] = timed(ctx, step)
        res[name] = (out.to_host(), grad.to_host())
    errs = {"mixed_vs_fp64": rel(res["mixed"][0], res["fp64"][0]),
            "mixed_vs_fp64_gradient": rel(res["mixed"][1], res["fp64"][1])}

    def cpu_grad(n):
        xn = tuple(c[:n] for c in x)
        orc.calculate(orc.MAGGRAD, m, mv, xn, field=field)
        orc.lqderivative(orc.MAGGRAD, m, mv, xn, rtf[:n], field=field)
    cpu = cpu_rate(cpu_grad, d.nsites, budget=4.0)
    row(4, "magnetic vertical gradient forward + adjoint", m, d, times, "maggrad", errs, cpu,
        passes=2)
    return rows


def cpu_model():
    """The host CPU's model name, family/model/stepping and clock (first processor in
    /proc/cpuinfo; virtual machines may report only a short model name), and the logical
    CPUs the process sees."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # the first processor only
                key, _, val = line.partition(":")
                fields[key.strip()] = val.strip()
    except OSError:
        pass
    if "model name" not in fields:
        import platform
        return platform.processor() or "unknown"
    extra = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    if "cpu MHz" in fields:
        extra.append(f"{float(fields['cpu MHz']):.0f} MHz")
    extra.append(f"{os.cpu_count()} logical CPUs visible")
    return f"{fields['model name']} ({', '.join(extra)})"


def cpu_baseline(model, data, budget_s, runs=5):
    """The oracle (the OpenMP restatement of the reference's OMP imps) on a bounded,
    fixed sample of sites: gz and FTG calculators separately, as jif3D runs them;
    Calculate (forward) and LQDerivative (matrix-free adjoint) timed separately; one
    warm-up, then the median of `runs` timed runs of each (steady clock).  The sample
    size is chosen so that the timed runs take about `budget_s` seconds; work per site
    is identical, so the rate scales linearly to all sites."""
    from oracle import oracle as orc
    mv = model.GetModelParameters()
    x, y, z = data.positions
    S, N = data.nsites, model.ncells
    rng = np.random.default_rng(11)
    order = rng.permutation(S)

    def fwd(idx):
        sites = (x[idx], y[idx], z[idx])
        t = time.perf_counter()
        g = orc.calculate(orc.SCALAR, model, mv, sites)
        tt = orc.calculate(orc.TENSOR, model, mv, sites)
        return time.perf_counter() - t, g, tt

    def adj(idx):
        sites = (x[idx], y[idx], z[idx])
        r1 = np.random.default_rng(12).normal(size=len(idx))
        r9 = np.random.default_rng(13).normal(size=9 * len(idx))
        t = time.perf_counter()
        orc.lqderivative(orc.SCALAR, model, mv, sites, r1)
        orc.lqderivative(orc.TENSOR, model, mv, sites, r9)
        return time.perf_counter() - t

    probe = order[:1]
    t1 = fwd(probe)[0] + adj(probe)  # also the warm-up of the thread pool
    n = int(max(1, min(S - 1, budget_s / (runs + 1) / max(t1, 1e-3))))
    idx = order[1:1 + n]
    fwd(idx)  # warm-up
    adj(idx)
    tf, ta = [], []
    for _ in range(runs):
        t, g, tt = fwd(idx)
        tf.append(t)
        ta.append(adj(idx))
    mf, ma = float(np.median(tf)), float(np.median(ta))
    pairs = len(idx) * N
    return {"value": pairs / (mf + ma), "unit": "site·cell pairs/s (forward + adjoint)",
            "cores": orc.num_threads(), "kind": "port", "cpu_model": cpu_model(),
            "forward_pairs_per_s": pairs / mf, "adjoint_pairs_per_s": pairs / ma,
            "statistic": f"median of {runs} timed runs after 1 warm-up, steady clock",
            "sample": f"{len(idx)} random sites x {N} cells (fixed seed), gz and FTG "
                      f"calculators separately; forward {mf:.2f} s and adjoint {ma:.2f} s "
                      f"per run (median)",
            "sites": idx, "gz": g, "ftg": tt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="both", choices=["mixed", "fp64", "both"])
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="quarter-size mesh for quick checks")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 512 MiB device-to-device copy (counter calibration)")
    ap.add_argument("--workload", default="all", choices=["all", "headline", "storedj", "configs"],
                    help="headline: gz+FTG forward+adjoint only; storedj: batched stored-J only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards "
              "by sites (forward) and cells (adjoint)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "storedj":
        print(json.dumps({"storedj": run_storedj(args)}), flush=True)
        return
    if args.workload == "configs":
        print(json.dumps({"configs": run_configs(args)}), flush=True)
        return

    from jif3d_amd import synthetic
    if args.small:
        model, data = synthetic.config4_gravity(n=64, nzc=32, nsite=32)
    else:
        model, data = synthetic.config4_gravity()
    S, N = data.nsites, model.ncells
    rho_c = (model.nx + 1) * (model.ny + 1) * (model.nz + 1) / N
    log(f"[bench] workload: {model.nx}x{model.ny}x{model.nz} cells, {S} sites")

    if args.calibrate_copy:
        from jif3d_amd import _native as nat
        cctx = nat.Context(0)
        a = cctx.alloc(64 * 1024 * 1024)  # 512 MiB of doubles
        b = cctx.alloc(64 * 1024 * 1024)
        for _ in range(2):
            cctx.copy_device(b, a)
        cctx.synchronize()
        a.free()
        b.free()
        log("[bench] calibration copy: 2 x 512 MiB device-to-device")
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {p: run_gpu(args, model, data, p) for p in precs}
    main_p = "mixed" if "mixed" in res else "fp64"
    r = res[main_p]

    # dominant kernel and its roofline: the FLOPs it executes per launch (hardware
    # counters, committed summary) over its average launch time measured here with HIP
    # events, against the vector rate the kernel can reach (FP64 vector, or FP32 without
    # packed instructions); the closed-form counting contract is reported beside it as an
    # effective rate, since the multipole far field needs far fewer operations per pair
    kern = max(r["kernels"].items(), key=lambda kv: kv[1][0])
    kname, (kms, klaunch) = kern
    avg_s = kms / klaunch * 1e-3
    fe = S * N * fe_per_pair(rho_c)  # forward-equivalent work of one launch
    is64 = "fp64" in kname
    peak = PEAK["fp64"] if is64 else FP32_UNPACKED
    fpp = profiled_flops_per_pair(kname)
    roof = {"bound": "valu", "kernel": kname, "avg_launch_ms": avg_s * 1e3}
    if fpp is not None:
        flops = fpp * S * N
        roof.update({"achieved": flops / avg_s / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": flops / avg_s / peak,
                     "flops_per_launch": flops,
                     "work": f"{fpp} FP32+FP64 FLOPs per site-cell pair (SQ_INSTS_VALU_FLOPS_* "
                             f"x 64 lanes per dispatch / pairs, profiles/"
                             f"gz_ftg_128x128x64_valu_counters.md) x {S * N} pairs",
                     "peak_name": ("FP64 vector" if is64 else
                                   "FP32 vector without packed-FP32 instructions (disabled in "
                                   "this build); the published 157.3 TFLOP/s counts packed FP32"),
                     "frac_of_published_fp32_peak": None if is64 else flops / avg_s / PEAK["fp32"]})
    roof["traffic"] = profiled_traffic(kname)
    roof["effective_fe_rate"] = {
        "value": fe / avg_s / 1e12, "unit": "TFLOP/s-equivalent",
        "work": f"{fe_per_pair(rho_c):.1f} flop-equivalents per pair (closed-form counting "
                f"contract) x {S * N} pairs per launch; not a roofline: the far field "
                f"replaces the closed form by a multipole"}
    line = {
        "metric": METRIC,
        "value": r["pairs_per_s"],
        "unit": "site·cell pairs/s (forward + adjoint per pair)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_p == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"gz+FTG forward+adjoint, {model.nx}x{model.ny}x{model.nz} cells "
                               f"(20 m), {int(np.sqrt(S))}x{int(np.sqrt(S))} sites at z=-30 m",
                   "cells": N, "sites": S, "components": 10, "precision": main_p},
        "roofline": roof,
    }
    if "fp64" in res and main_p != "fp64":
        line["fp64_pairs_per_s"] = res["fp64"]["pairs_per_s"]
        line["speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / r["ms_per_step"]
        ref = res["fp64"]["out"].reshape(-1, 10)
        mix = r["out"].reshape(-1, 10)
        comp_err = [float(np.max(np.abs(mix[:, c] - ref[:, c])) / np.max(np.abs(ref[:, c])))
                    for c in range(10)]
        gz_pt = float(np.max(np.abs(mix[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])))
        # pointwise, per value: |mixed - FP64| / |FP64| over the values at least 1 % of their
        # component's survey maximum (tensor components cross zero inside the survey)
        big = np.abs(ref) >= 0.01 * np.max(np.abs(ref), axis=0)
        pt = np.abs(mix - ref)[big] / np.abs(ref)[big]
        gref, gmix = res["fp64"]["grad"], r["grad"]
        line["error_vs_fp64"] = {
            "max_rel_per_component_inf_norm": max(comp_err),
            "per_component_inf_norm": dict(zip(
                ["gz", "Uxx", "Uxy", "Uxz", "Uyx", "Uyy", "Uyz", "Uzx", "Uzy", "Uzz"], comp_err)),
            "max_pointwise_rel_above_1pct_of_max": float(np.max(pt)),
            "gz_max_pointwise_rel": gz_pt,
            "gradient_rel_inf_norm": float(np.max(np.abs(gmix - gref)) / np.max(np.abs(gref)))}
    line["kernels_ms_per_step"] = {k: v[0] / args.steps for k, v in r["kernels"].items()}
    if args.workload == "all":
        line["storedj"] = run_storedj(args)
    if not args.no_cpu:
        cb = cpu_baseline(model, data, args.cpu_budget)
        idx = cb.pop("sites")
        g, t9 = cb.pop("gz"), cb.pop("ftg")
        ref64 = res["fp64"]["out"].reshape(-1, 10) if "fp64" in res else None
        if ref64 is not None:
            e_gz = float(np.max(np.abs(ref64[idx, 0] - g) / np.abs(g)))
            t9 = t9.reshape(-1, 9)
            e_t = float(np.max(np.abs(ref64[idx, 1:] - t9)) / np.max(np.abs(t9)))
            cb["gpu_fp64_vs_oracle"] = {"gz_max_rel": e_gz, "ftg_rel_inf_norm": e_t}
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
