#!/usr/bin/env python3
"""Benchmark of the GP fit+predict path on one MI355X.

A step is one fit+predict request at the plain-FP64 scale configuration:
n_train = 16384 (32 tiles of 512), n_reg = 8, n_test = 4096, theta = (1, 1, 0.1),
seeded mass-spring-damper data. The request assembles K, factors it, solves
alpha = K^-1 y and returns the predictive mean and latent variance at the
test points (the reference's predict_with_uncertainty). Inputs are placed in
HBM before the timed region and outputs stay there, so `value` excludes PCIe.

Also reported on the same line:
  * p50 / p99 request latency over the timed steps;
  * the n = 32768 FP64 Cholesky rate (n^3/3 flops over the factorisation's
    wall time, n_reg = 16, noise 0.05);
  * `roofline` for the dominant kernel (the Cholesky trailing update on FP64
    MFMA), from HIP events bracketing each of its launches in the timed
    region, against the 78.6 TFLOP/s FP64 matrix peak;
  * `cpu_baseline`: the CPU oracle (a restatement of the reference's HPX +
    BLAS path, oneMKL sequential per tile on a host pool) timed on one full
    request on this host's cores, and the GPU result's error against it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fit+predict req/s & p50/p99 latency at ≤1e-8 rel err; n=32k Cholesky TFLOP/s"
FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

N_TILES, TILE, N_REG, N_TEST = 32, 512, 8, 4096
THETA = (1.0, 1.0, 0.1)


def outer_block(n_pad: int) -> int:
    env = os.environ.get("GPRAT_CHOLESKY_BLOCK")
    if env and int(env) >= 128 and int(env) % 128 == 0:
        return int(env)
    return 256 if n_pad <= 4096 else 512


def trailing_update_flops(n_pad: int):
    """Algorithmic flops of each trailing-update launch, in launch order:
    2 * K * (number of lower-triangle elements updated)."""
    bo = outer_block(n_pad)
    out = []
    p0 = 0
    while p0 < n_pad:
        pe = min(n_pad, p0 + bo)
        if pe >= n_pad:
            break
        q1 = min(n_pad, pe + bo)
        K = pe - p0
        M, N = n_pad - pe, q1 - pe
        out.append(2.0 * K * (N * M - N * (N - 1) / 2))
        if q1 < n_pad:
            S = n_pad - q1
            out.append(2.0 * K * (S * (S + 1) / 2))
        p0 = pe
    return out


def percentile(values, q):
    return float(np.percentile(np.asarray(values), q))


def cpu_baseline(x, y, z, gpu_mean, gpu_var):
    sys.path.insert(0, os.path.join(ROOT, "oracle", "_build"))
    import gprat_oracle
    threads = min(os.cpu_count() or 1, 16)
    gprat_oracle.set_threads(threads)
    mt, mb = gprat_oracle.compute_test_tiles(N_TEST, N_TILES, TILE)
    gp = gprat_oracle.GP(list(x), list(y), N_TILES, TILE, N_REG, list(THETA), [True, True, True])
    t0 = time.perf_counter()
    mean, var = gp.predict_with_uncertainty(list(z), mt, mb)
    dt = time.perf_counter() - t0
    mean, var = np.asarray(mean), np.asarray(var)
    err_mean = float(np.max(np.abs(gpu_mean - mean)) / np.max(np.abs(mean)))
    err_var = float(np.max(np.abs(gpu_var - var)) / np.max(np.abs(var)))
    return {
        "value": 1.0 / dt,
        "unit": "req/s",
        "cores": threads,
        "kind": "port",
        "sample": f"1 full fit+predict request (n={N_TILES * TILE}, m={N_TEST}) on the oracle, "
                  f"{gprat_oracle.blas_name()} sequential per tile, {threads} pool threads; {dt:.2f} s",
    }, {"mean": err_mean, "var": err_var}


def cholesky_rate(capi, ctx, n=32768, repeats=2):
    """FP64 Cholesky TFLOP/s at n (n_reg 16, noise 0.05), wall time of the
    factorisation from HIP events around it."""
    from gprat_amd import datagen
    x, y, _ = datagen.gp_problem(n, 1, 16)
    dx, dy = ctx.device_array(x), ctx.device_array(y)
    theta = (1.0, 1.0, 0.05)
    capi.Model(ctx, dx, dy, n, 16, theta).release()  # warm-up (allocations, code objects)
    best = None
    times = []
    for _ in range(repeats):
        ctx.set_profiling(True)
        m = capi.Model(ctx, dx, dy, n, 16, theta)
        cnt, ms = ctx.kernel_stats("cholesky")
        ctx.set_profiling(False)
        m.release()
        times.append(ms)
        best = ms if best is None else min(best, ms)
    dx.free()
    dy.free()
    flops = n ** 3 / 3.0
    return {"n": n, "ms_best": best, "ms_mean": float(np.mean(times)), "tflops": flops / (best * 1e-3) / 1e12}


def mixed_leg(capi, ctx, n=32768, m=8192, r=16, noise=0.05, repeats=2):
    """Mixed-precision configuration: n=32768 (32 tiles of 1024), n_reg 16,
    m=8192, noise 0.05. Times one fit+predict request in FP64 and in mixed
    precision (FP32 matrix-core trailing updates + FP64 refinement of
    alpha) and reports the mixed result's error against the FP64 one."""
    from gprat_amd import datagen
    x, y, z = datagen.gp_problem(n, m, r)
    theta = (1.0, 1.0, noise)
    dx, dy, dz = ctx.device_array(x), ctx.device_array(y), ctx.device_array(z)
    out = {}
    res = {}
    for name, prec in (("fp64", capi.PRECISION_FP64), ("mixed", capi.PRECISION_MIXED)):
        times = []
        for rep in range(repeats + 1):
            ctx.synchronize()
            t0 = time.perf_counter()
            mdl = capi.Model(ctx, dx, dy, n, r, theta, prec)
            mean, var = mdl.predict(dz, m)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            if rep > 0:
                times.append(dt)
            if rep == repeats:
                res[name] = (mean, var, mdl.loss(), mdl.refinement())
            mdl.release()
        out[name + "_ms"] = 1e3 * min(times)
    m64, v64, l64, _ = res["fp64"]
    mmx, vmx, lmx, (it, last) = res["mixed"]
    out["speedup"] = out["fp64_ms"] / out["mixed_ms"]
    out["rel_err_vs_fp64"] = {
        "mean": float(np.max(np.abs(mmx - m64)) / np.max(np.abs(m64))),
        "var": float(np.max(np.abs(vmx - v64)) / np.max(np.abs(v64))),
        "loss": abs(lmx - l64) / abs(l64),
    }
    out["refinement_steps"] = it
    out["workload"] = f"fit+predict n={n} (32 tiles of 1024), n_reg={r}, m={m}, noise={noise}"
    for a in (dx, dy, dz):
        a.free()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--chol-n", type=int, default=32768, help="size of the Cholesky-rate leg (0 skips it)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this backend runs on one GPU; the GP path has no multi-GPU mode in this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from gprat_amd import capi, datagen

    n = N_TILES * TILE
    x, y, z = datagen.gp_problem(n, N_TEST, N_REG)
    ctx = capi.Context(0, 1)
    dx, dy, dz = ctx.device_array(x), ctx.device_array(y), ctx.device_array(z)
    dmean, dvar = ctx.device_empty(N_TEST), ctx.device_empty(N_TEST)

    def step():
        model = capi.Model(ctx, dx, dy, n, N_REG, THETA)
        model.predict(dz, N_TEST, variance=True, out=(dmean, dvar))
        model.release()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    ctx.set_profiling(True)
    lat = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        step()
        lat.append(time.perf_counter() - s)
    ctx.synchronize()
    total = time.perf_counter() - t0
    fams = {f: ctx.kernel_stats(f) for f in
            ("syrk_trailing", "panel", "cholesky", "assemble", "solve", "trsm", "mean")}
    ctx.set_profiling(False)
    gpu_mean, gpu_var = dmean.numpy(), dvar.numpy()

    per_launch = trailing_update_flops(((n + 127) // 128) * 128)
    launches, syrk_ms = fams["syrk_trailing"]
    expect = len(per_launch) * args.steps
    syrk_flops = sum(per_launch) * args.steps
    achieved = syrk_flops / (syrk_ms * 1e-3) / 1e12 if syrk_ms > 0 else None
    roofline = {
        "kernel": "k_gemm_f64<false,true,0> (Cholesky trailing update, FP64 MFMA)",
        "bound": "mfma",
        "achieved": achieved,
        "peak": FP64_MFMA_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": (achieved / FP64_MFMA_PEAK_TFLOPS) if achieved else None,
        "traffic": None,
        "launches": launches,
        "launches_expected": expect,
        "avg_launch_ms": syrk_ms / launches if launches else None,
        "flops_per_launch_avg": syrk_flops / launches if launches else None,
    }
    chol_ms = fams["cholesky"][1] / max(fams["cholesky"][0], 1)

    chol = None
    if args.chol_n > 0:
        chol = cholesky_rate(capi, ctx, args.chol_n)
    mixed = None if args.no_mixed else mixed_leg(capi, ctx)

    result = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "req/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mass-spring-damper series, SplitMix64)",
        "config": {
            "workload": f"fit+predict request: n_train={n} ({N_TILES} tiles of {TILE}), n_reg={N_REG}, "
                        f"n_test={N_TEST}, theta={THETA}, FP64, inputs resident in HBM",
            "precision": "fp64",
        },
        "p50_ms": percentile(lat, 50) * 1e3,
        "p99_ms": percentile(lat, 99) * 1e3,
        "cholesky_n16384_ms": chol_ms,
        "cholesky_n16384_tflops": (n ** 3 / 3.0) / (chol_ms * 1e-3) / 1e12 if chol_ms > 0 else None,
        "cholesky_n32k": chol,
        "mixed_precision_n32k": mixed,
        "phase_ms_per_step": {k: v[1] / args.steps for k, v in fams.items()},
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        base, err = cpu_baseline(x, y, z, gpu_mean, gpu_var)
        result["cpu_baseline"] = base
        result["rel_err_vs_oracle"] = err
    else:
        result["cpu_baseline"] = None
    for a in (dx, dy, dz, dmean, dvar):
        a.free()
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
