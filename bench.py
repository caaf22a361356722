"""Benchmark: gene x cell fits per second of the per-gene NB-GLM fit on one MI355X.

One step = the whole fitting path over a synthetic count matrix already
resident in HBM: init pass (alpha0, beta0), IRLS, overdispersion, bread (+
clustered meat when the workload has patients) and the Wald test with BH.
Default workload: G=20000 genes x n=500000 cells int32 counts, p=8 design
(intercept, 3-level condition, five covariates), MOM overdispersion, fitted
with the mixed-precision IRLS; the same counts are then fitted with the FP64
path, which gives the speed-up and the accuracy check (beta, log10 p, BH
calls) of the mixed path against FP64.

Prints one JSON line with the throughput, a roofline entry for the dominant
kernel (the IRLS accumulate pass: algorithmic bytes = count element size x n
x active genes per launch, over its average launch duration from HIP events
on the library's stream) and a CPU baseline: the FP64 oracle timed on a gene
subsample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "gene×cell fits/s: mixed-precision vs FP64 GPU path at ≤1e-4 β/p rel err"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_FLOPS = 2.5e15


def hbm_copy_gbps(dev, nbytes=4 << 30, reps=10):
    """HBM bandwidth of a plain device copy (torch's copy kernel) of nbytes:
    bytes read + written per copy over its event-timed duration."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.int32, device=dev)
    b = torch.empty_like(a)
    a.fill_(1)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    gbps = 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbps


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries, by workload
TRAFFIC_FILES = {"C3": "traffic_irls_C3.json", "C4": "traffic_C4_mixed_clustered.json"}


def measured_traffic(config, kdesc, alg_bytes_per_launch):
    """HBM bytes per launch of the kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement of the same workload (traffic per algorithmic byte), or
    None."""
    if config not in TRAFFIC_FILES:
        return None
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILES[config])
    try:
        with open(path) as f:
            meas = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    name = kdesc.split(" ")[0]
    if name not in meas:
        return None
    return meas[name]["traffic_per_algorithmic_byte"] * alg_bytes_per_launch


def host_cpu():
    """Host facts recorded with the CPU baseline: logical CPUs the OS reports,
    the CPUs this process may run on, and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity": allowed, "model": model}


def rel_err(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    e = np.abs(a - b) / np.maximum(np.abs(b), floor)
    e[np.isnan(a) & np.isnan(b)] = 0.0
    return float(np.max(e)) if e.size else 0.0


def cpu_baseline(w, Yd, fits, n_threads, max_genes=None):
    """The FP64 oracle on an evenly strided gene subsample (about 10 s of host
    work on n_threads threads, plus a 1-thread run on an eighth of it), timed
    on the host; the oracle's fit of the sample also checks every GPU fit in
    ``fits`` ({precision: (fit, test_de result)}) gene by gene."""
    import oracle
    cost = w.n * w.p * w.p * (4 if w.od == "MLE" else 1)
    target = max_genes or int(np.clip(round(4096 * 6.4e6 / cost), 32, 4096))
    stride = math.ceil(w.G / target)
    genes = np.arange(0, w.G, stride)
    Ys = Yd[genes].cpu().numpy().astype(np.float64)
    od = {"MOM": oracle.OD_MOM, "MLE": oracle.OD_MLE}[w.od]
    ends = w.cluster_ends
    contrast = np.zeros(w.p)
    contrast[1] = 1.0
    log(f"[bench] cpu baseline: {len(genes)} genes x {w.n} cells, {n_threads} threads")
    t0 = time.perf_counter()
    ref = oracle.cpu_fit(Ys, w.X, w.offset, clusters_ends=ends, od_mode=od,
                         n_threads=n_threads)
    rde = oracle.de_test(ref["beta"], ref["bread"], ref.get("sandwich"), w.n, contrast)
    t = time.perf_counter() - t0
    g1 = genes[::8]
    log(f"[bench] cpu baseline: {len(g1)} genes x {w.n} cells, 1 thread")
    t1 = time.perf_counter()
    ref1 = oracle.cpu_fit(Ys[::8], w.X, w.offset, clusters_ends=ends, od_mode=od, n_threads=1)
    oracle.de_test(ref1["beta"], ref1["bread"], ref1.get("sandwich"), w.n, contrast)
    t1 = time.perf_counter() - t1
    check = {"genes": int(len(genes))}
    for prec, (fit, de) in fits.items():
        beta = fit["beta"].cpu().numpy()[genes]
        it = fit["iterations"]["beta_iters"].cpu().numpy()[genes]
        th = fit["overdispersion"].cpu().numpy()[genes]
        pv = np.asarray(de["pval"])[genes]
        lp = np.asarray(de["log_pval"])[genes]
        bread = fit["beta_sandwiches_null"].cpu().numpy()[genes]
        rel_th = np.abs(th - ref["theta"]) / np.maximum(np.abs(ref["theta"]), 1e-8)
        worst = int(np.nanargmax(rel_th))
        se = np.sqrt(np.abs(np.diagonal(bread, axis1=1, axis2=2)))
        se_ref = np.sqrt(np.abs(np.diagonal(ref["bread"], axis1=1, axis2=2)))
        check[prec] = {
            "beta_max_rel_err_floor_1e-6": rel_err(beta, ref["beta"], 1e-6),
            "beta_max_rel_err_floor_1e-3": rel_err(beta, ref["beta"], 1e-3),
            "theta_max_rel_err": float(np.nanmax(rel_th)),
            "theta_worst_gene": {"gene": int(genes[worst]), "gpu": float(th[worst]),
                                 "oracle": float(ref["theta"][worst])},
            "se_max_rel_err": rel_err(se, se_ref, 1e-300),
            "pval_max_rel_err": rel_err(pv, rde["pval"], 1e-300),
            "log_pval_max_rel_err": rel_err(lp, rde["log_pval"], 1e-300),
            "iter_mismatch": int(np.sum(it != ref["beta_iters"]))}
    hc = host_cpu()
    return {"value": len(genes) * w.n / t, "unit": "gene*cell fits/s", "cores": n_threads,
            "kind": "port",
            "sample": f"FP64 oracle cpu_fit + test_de on every {stride}th gene "
                      f"({len(genes)} genes) x all {w.n} cells of the same workload, "
                      f"{t:.1f} s on {n_threads} threads (the host's CPU share for one GPU; "
                      f"nproc {hc['nproc']}, {hc['affinity']} CPUs in the affinity mask, "
                      f"{hc['model']})",
            "one_thread": {"value": len(g1) * w.n / t1, "genes": int(len(g1)),
                           "seconds": t1},
            "host": hc}, check


def compare(dv, mix, f64, de_mix, de_64):
    """Mixed against FP64 on the same counts: the accuracy side of the metric."""
    def rel(a, b, floor=1e-6):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        e = np.abs(a - b) / np.maximum(np.abs(b), floor)
        e[np.isnan(a) & np.isnan(b)] = 0.0
        return float(np.max(e))
    cpu = lambda t: t.cpu().numpy() if hasattr(t, "cpu") else np.asarray(t)
    lp_m = cpu(de_mix["log_pval"]) / math.log(10)
    lp_f = cpu(de_64["log_pval"]) / math.log(10)
    calls_m = cpu(de_mix["adj_pval"]) < 0.05
    calls_f = cpu(de_64["adj_pval"]) < 0.05
    it_m = cpu(mix["iterations"]["beta_iters"])
    it_f = cpu(f64["iterations"]["beta_iters"])
    def worst(a, b, floor):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        e = np.abs(a - b) / np.maximum(np.abs(b), floor)
        e[np.isnan(a) & np.isnan(b)] = 0.0
        i = int(np.nanargmax(e))
        return {"index": i, "mixed": float(a.flat[i]), "fp64": float(b.flat[i])}
    sig = np.abs(lp_f) >= 1.0  # p <= 0.1
    th_m, th_f = cpu(mix["overdispersion"]), cpu(f64["overdispersion"])
    se = lambda f: np.sqrt(np.abs(np.diagonal(cpu(f["beta_sandwiches_null"]), axis1=1, axis2=2)))
    return {"definition": "relative error |mixed - fp64| / max(|fp64|, floor); the 1e-4 "
                          "target holds for beta at floor 1e-3 and log10 p at floor 1, not "
                          "at floor 1e-6 (beta_max_rel_err_floor_1e-6, log10p_max_rel_err)",
            "se_max_rel_err": rel(se(mix), se(f64), 1e-300),"beta_max_rel_err": rel(cpu(mix["beta"]), cpu(f64["beta"]), 1e-3),
            "log10p_max_rel_err_floor_1": rel(lp_m, lp_f, 1.0),
            "log10p_max_rel_err_p_le_0.1": rel(lp_m[sig], lp_f[sig]) if sig.any() else 0.0,
            "log10p_worst": worst(lp_m, lp_f, 1e-6),
            "theta_max_rel_err_floor_1e-3": rel(th_m, th_f, 1e-3),
            "theta_worst": worst(th_m, th_f, 1e-8),
            "beta_max_rel_err_floor_1e-6": rel(cpu(mix["beta"]), cpu(f64["beta"]), 1e-6),
            "beta_max_abs_err": float(np.nanmax(np.abs(cpu(mix["beta"]) - cpu(f64["beta"])))),
            "log10p_max_rel_err": rel(lp_m, lp_f),
            "theta_max_rel_err": rel(cpu(mix["overdispersion"]), cpu(f64["overdispersion"]),
                                     1e-8),
            "bh_calls_fp64": int(calls_f.sum()),
            "bh_call_mismatch": int(np.sum(calls_m != calls_f)),
            "iter_mismatch": int(np.sum(it_m != it_f))}


def bench_many(args):
    """C5: many independent problems in one devil_fit call (mixed precision,
    autotuned kernels) against the FP64 path run problem by problem."""
    import torch

    import devil_amd
    from devil_amd import synth

    dev = torch.device("cuda", 0)
    ws = synth.make_problems("C5", G=args.genes)
    log(f"[bench] generating C5: {len(ws)} problems, G={ws[0].G}, "
        f"sum n={sum(w.n for w in ws)} on the GPU")
    Ys = [synth.counts_device(w, device=dev) for w in ws]
    torch.cuda.synchronize()
    ctx = devil_amd.context(0)
    ctx.set_timing(True)
    if args.concurrency:
        from devil_amd import _lib
        ctx.set_tuning(_lib.TUNE_CONCURRENCY, args.concurrency)
    if args.pad_copy >= 0:
        from devil_amd import _lib
        ctx.set_tuning(_lib.TUNE_PAD_COPY, args.pad_copy)
    # every input resident in HBM before timing (design column-major, as read)
    jobs = [dict(x=Y, design_matrix=torch.as_tensor(np.ascontiguousarray(w.X.T), device=dev).t(),
                 size_factors=torch.as_tensor(np.asarray(w.sf, dtype=np.float64), device=dev))
            for w, Y in zip(ws, Ys)]
    units = sum(w.G * w.n for w in ws)

    def many():
        return devil_amd.fit_devil_many(jobs, precision="mixed")

    def one_by_one():
        return [devil_amd.fit_devil(**j, precision="fp64") for j in jobs]

    for i in range(args.warmup):
        many()
        torch.cuda.synchronize()
        log(f"[bench] warmup {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fits = many()
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    kall = ctx.kernel_stats()
    ks = kall["irls_mixed"]
    ctx_fp64_launches = kall["irls_fp64"][0]
    info = ctx.fit_info()
    log(f"[bench] mixed, {len(jobs)} problems per call: {1e3 * T / args.steps:.1f} ms/step")
    out = {
        "metric": METRIC, "value": units * args.steps / T, "unit": "gene*cell fits/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * T / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"C5: {len(ws)} independent problems, G={ws[0].G} shared genes, "
                               f"n_j log-uniform in [500, 2e5] (sum {sum(w.n for w in ws)}), "
                               "p_j = 2 + (j mod 9), MOM, int32 counts; one devil_fit call "
                               "(mixed precision, autotuned kernel and chunk size)",
                   "G": ws[0].G, "n_total": sum(w.n for w in ws), "problems": len(ws),
                   "precision": "mixed"},
        "roofline": {"bound": "hbm", "kernel": "mixed IRLS accumulate (all problems)",
                     "achieved": ks[2] / (ks[1] * 1e-3) / 1e9 if ks[1] > 0 else 0.0,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (ks[2] / (ks[1] * 1e-3) / 1e9 / HBM_PEAK_GBS) if ks[1] > 0 else 0.0,
                     "traffic": None, "launches_last_step": ks[0]},
        "fit_info": info,
    }
    # the same call with joint IRLS rounds (one accumulate and one update launch
    # per round for each unit of problems of one design width)
    from devil_amd import _lib
    ctx.set_tuning(_lib.TUNE_JOINT, 1)
    try:
        many()  # tunes the joint rounds' chunk sizes
        torch.cuda.synchronize()
        tj = time.perf_counter()
        for _ in range(args.steps):
            many()
        torch.cuda.synchronize()
        TJ = (time.perf_counter() - tj) / args.steps
        kj = ctx.kernel_stats()
        out["joint_rounds"] = {
            "ms_per_step": 1e3 * TJ, "value": units / TJ, "problems_joined": ctx.fit_info()["joined"],
            "launches_mixed_accumulate": kj["irls_mixed"][0],
            "launches_fp64_accumulate": kj["irls_fp64"][0],
            "launches_own_rounds": {"mixed": ks[0], "fp64": ctx_fp64_launches}}
    finally:
        ctx.set_tuning(_lib.TUNE_JOINT, 0)
    log(f"[bench] joint rounds: {1e3 * TJ:.1f} ms/step")
    if args.fp64_steps > 0:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.fp64_steps):
            ref = one_by_one()
        torch.cuda.synchronize()
        T64 = (time.perf_counter() - t1) / args.fp64_steps
        log(f"[bench] fp64 problem by problem: {1e3 * T64:.1f} ms")
        eb = 0.0
        it_mis = 0
        for f, r in zip(fits, ref):
            a, b = f["beta"].cpu().numpy(), r["beta"].cpu().numpy()
            eb = max(eb, float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3))))
            it_mis += int((f["iterations"]["beta_iters"] != r["iterations"]["beta_iters"]).sum())
        out["fp64"] = {"ms_per_step": 1e3 * T64, "value": units / T64,
                       "speedup_mixed_vs_fp64": T64 / (T / args.steps)}
        out["mixed_vs_fp64"] = {"beta_max_rel_err": eb, "iter_mismatch": it_mis}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C3")
    ap.add_argument("--genes", type=int, default=None)
    ap.add_argument("--cells", type=int, default=None)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--fp64-steps", type=int, default=1,
                    help="timed FP64 steps on the same counts (0: no comparison)")
    ap.add_argument("--cpu-genes", type=int, default=None)
    ap.add_argument("--mixed-kernel", default=None, choices=["f32", "bf16", "bf16_ws", "bf16_ws2"],
                    help="mixed IRLS kernel (default: the library's choice)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--chunk-cells", type=int, default=0,
                    help="cells per chunk of the tile passes (0: autotuned)")
    ap.add_argument("--pad-copy", type=int, default=-1, choices=[-1, 0, 1],
                    help="mixed fits of misaligned device rows: 1 read a padded copy, 0 read in "
                         "place (-1: library default)")
    ap.add_argument("--concurrency", type=int, default=0,
                    help="C5: host threads / streams of the many-problem call (0: library "
                         "default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the fit shards by genes with no communication; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    if args.config == "C5":
        return bench_many(args)

    import torch

    import devil_amd
    from devil_amd import synth

    dev = torch.device("cuda", 0)
    w = synth.make_workload(args.config, G=args.genes, n=args.cells)
    log(f"[bench] generating {args.config}: G={w.G} n={w.n} p={w.p} od={w.od} on the GPU")
    t0 = time.perf_counter()
    Yd = synth.counts_device(w, device=dev)
    torch.cuda.synchronize()
    zero_frac = float((Yd == 0).float().mean().item())
    log(f"[bench] counts ready in {time.perf_counter() - t0:.1f} s, zero fraction {zero_frac:.3f}")
    contrast = np.zeros(w.p)
    contrast[1] = 1.0
    hbm_copy = hbm_copy_gbps(dev)
    log(f"[bench] device copy: {hbm_copy:.0f} GB/s")
    ctx = devil_amd.context(0)
    ctx.set_timing(True)
    if args.mixed_kernel:
        ctx.set_mixed_kernel(args.mixed_kernel)
    if args.chunk_cells:
        from devil_amd import _lib
        ctx.set_tuning(_lib.TUNE_CHUNK_CELLS, args.chunk_cells)
    if args.pad_copy >= 0:
        from devil_amd import _lib
        ctx.set_tuning(_lib.TUNE_PAD_COPY, args.pad_copy)

    # every input resident in HBM before timing: counts, the design matrix
    # (column-major, as the library reads it) and the size factors
    Xdev = torch.as_tensor(np.ascontiguousarray(w.X.T), device=dev).t()
    sfdev = torch.as_tensor(np.asarray(w.sf, dtype=np.float64), device=dev)

    def step(prec):
        fit = devil_amd.fit_devil(Yd, Xdev, clusters=w.clusters, overdispersion=w.od,
                                  size_factors=sfdev, precision=prec)
        de = devil_amd.test_de(fit, contrast)
        return fit, de

    def timed(prec, steps):
        stage_sum = {k: 0.0 for k in devil_amd._lib.STAGES}
        kern = {k: [0, 0.0, 0.0] for k in devil_amd._lib.KERNELS}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            fit, de = step(prec)
            for k, v in ctx.stage_times().items():
                stage_sum[k] += v
            for k, v in ctx.kernel_stats().items():
                for j in range(3):
                    kern[k][j] += v[j]
        torch.cuda.synchronize()
        T = time.perf_counter() - t0
        log(f"[bench] {prec}: {steps} steps, {1e3 * T / steps:.1f} ms/step")
        return T, {k: v / steps for k, v in stage_sum.items()}, kern, fit, de

    for i in range(args.warmup):
        t1 = time.perf_counter()
        step(args.precision)
        torch.cuda.synchronize()
        log(f"[bench] warmup {i}: {time.perf_counter() - t1:.3f} s")
    T, stages, kern, fit, de = timed(args.precision, args.steps)
    info = ctx.fit_info()
    ms_step = 1e3 * T / args.steps
    iters = fit["iterations"]["beta_iters"].cpu().numpy().astype(np.int64)
    th_it = fit["theta_passes"].cpu().numpy().astype(np.int64)
    sum_it = int(iters.sum())
    kname = "irls_mixed" if args.precision == "mixed" else "irls_fp64"
    nl, kms, kbytes = kern[kname]
    avg_ms = kms / max(nl, 1)
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    # whole-fit algorithmic bytes: init + IRLS + dispersion passes +
    # bread/meat, one read of 4 bytes per count per pass
    disp_passes = int(th_it.sum()) if w.od == "MLE" else w.G
    fit_bytes = 4.0 * w.n * (w.G + sum_it + disp_passes + w.G)
    # algorithmic MFMA work of the IRLS contractions
    fit_flops = 2.0 * w.n * sum_it * (w.p * (w.p + 1) / 2 + 2 * w.p)
    fit_ms = sum(stages[k] for k in ("init", "irls", "dispersion", "bread_meat"))
    value = w.G * w.n * args.steps / T
    if kname == "irls_mixed" and info.get("mixed_kernel") == "bf16":
        kdesc = f"k_irls_accum_bf16<{w.p}> (bf16 MFMA, FP64 partials)"
    elif kname == "irls_mixed" and info.get("mixed_kernel") == "bf16_ws":
        kdesc = f"k_irls_accum_ws<{w.p}> (bf16 MFMA, LDS-staged operands)"
    elif kname == "irls_mixed" and info.get("mixed_kernel") == "bf16_ws2":
        kdesc = f"k_irls_accum_ws2<{w.p}> (bf16 MFMA, LDS-staged operands, two tiles per wave)"
    elif kname == "irls_mixed":
        kdesc = f"k_irls_accum_mixed<{w.p}> (FP32 MFMA, FP64 partials)"
    else:
        kdesc = f"k_irls_accum<{w.p}> (FP64 MFMA)"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "gene*cell fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: G={w.G} genes x n={w.n} cells int32 NB counts "
                               f"(zero fraction {zero_frac:.2f}), p={w.p}, {w.od} "
                               "overdispersion, bread"
                               + (", clustered sandwich" if w.clusters is not None else "")
                               + ", Wald test + BH",
                   "G": w.G, "n": w.n, "p": w.p, "precision": args.precision,
                   "mixed_kernel": args.mixed_kernel or "library default",
                   "overdispersion": w.od},
        "roofline": {"bound": "hbm", "kernel": kdesc,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.config, kdesc, kbytes / max(nl, 1)),
                     "launches_per_step": nl / args.steps,
                     "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": kbytes / max(nl, 1)},
        "stages_ms": stages,
        "fit_info": info,
        "fit_alg_GBps": fit_bytes / (fit_ms * 1e-3) / 1e9,
        "fit_alg_frac_of_copy": fit_bytes / (fit_ms * 1e-3) / 1e9 / hbm_copy,
        "hbm_copy_GBps": hbm_copy,
        "mfma_util": {"flops_alg_per_step": fit_flops,
                      "of_fit": fit_flops / (ms_step * 1e-3) / BF16_PEAK_FLOPS,
                      "of_irls_stage": fit_flops / (stages["irls"] * 1e-3) / BF16_PEAK_FLOPS
                      if stages["irls"] > 0 else 0.0,
                      "peak": "bf16 dense MFMA 2.5 PFLOP/s"},
        "sum_beta_iters": sum_it,
        "sum_theta_passes": int(th_it.sum()),
        "max_beta_iters": int(iters.max()),
    }
    nm, mms, mbytes = kern["mle"]
    if nm > 0:
        out["roofline_mle"] = {"bound": "hbm", "kernel": f"k_mle_accum<{w.p}> (FP64 MFMA)",
                               "achieved": mbytes / (mms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": mbytes / (mms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "launches_per_step": nm / args.steps,
                               "avg_launch_ms": mms / nm,
                               "algorithmic_bytes_per_launch": mbytes / nm}
    if args.precision == "mixed" and args.fp64_steps > 0:
        T64, st64, kern64, f64, de64 = timed("fp64", args.fp64_steps)
        ms64 = 1e3 * T64 / args.fp64_steps
        n64, kms64, kb64 = kern64["irls_fp64"]
        out["fp64"] = {"ms_per_step": ms64, "value": w.G * w.n / (ms64 * 1e-3),
                       "speedup_mixed_vs_fp64": ms64 / ms_step,
                       "stages_ms": st64,
                       "irls_kernel_GBps": kb64 / (kms64 * 1e-3) / 1e9 if kms64 > 0 else 0.0}
        out["mixed_vs_fp64"] = compare(devil_amd, fit, f64, de, de64)
    if not args.no_cpu_baseline:
        # the host's CPU share for one GPU is 16 threads
        n_threads = max(1, min(16, host_cpu()["affinity"]))
        fits = {args.precision: (fit, de)}
        if args.precision == "mixed" and args.fp64_steps > 0:
            fits["fp64"] = (f64, de64)
        cb, check = cpu_baseline(w, Yd, fits, n_threads, args.cpu_genes)
        out["cpu_baseline"] = cb
        out["cpu_check"] = check
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
