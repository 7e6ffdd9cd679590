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
FP32_MFMA_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

N_TILES, TILE, N_REG, N_TEST = 32, 512, 8, 4096
THETA = (1.0, 1.0, 0.1)


def outer_block(n_pad: int) -> int:
    env = os.environ.get("GPRAT_CHOLESKY_BLOCK")
    if env and int(env) >= 128 and int(env) % 128 == 0:
        return int(env)
    return 256 if n_pad <= 4096 else (512 if n_pad < 12288 else 768)


def trailing_update_flops(n_pad: int):
    """Algorithmic flops of each trailing-update launch, in launch order:
    2 * K * (number of elements updated). Step k's launch applies panel k
    (K = its width) to the rows below the next diagonal block: the next
    panel's columns whole, the columns after it lower (the next diagonal
    block itself is updated on the panel stream)."""
    bo = outer_block(n_pad)
    out = []
    p0 = 0
    while p0 < n_pad:
        pe = min(n_pad, p0 + bo)
        if pe >= n_pad:
            break
        q1 = min(n_pad, pe + bo)
        if q1 < n_pad:
            K = pe - p0
            S = n_pad - q1
            out.append(2.0 * K * (S * (q1 - pe) + S * (S + 1) / 2))
        p0 = pe
    return out


def measured_traffic():
    """HBM bytes per launch of the trailing update at this workload, from the
    committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE per
    dispatch, each scaled by what the counter reads for a 1 GiB copy);
    None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "trailing_update_n16384_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


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
    """Mixed-precision configuration: n in {32768, 65536} (tiles of 1024),
    n_reg 16, m=8192, noise 0.05. Times one fit+predict request in FP64 and
    in mixed precision (FP32 matrix-core trailing updates + FP64 refinement
    of alpha, variance solve on FP32 matrix cores) and reports the mixed
    result's error against the FP64 one."""
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
    out["workload"] = f"fit+predict n={n} ({n // 1024} tiles of 1024), n_reg={r}, m={m}, noise={noise}"
    for a in (dx, dy, dz):
        a.free()
    return out


def batched_leg(capi, ctx, count=2048, seed=4):
    """Batched many-GP configuration: 2048 independent GPs, n log-uniform in
    [64, 2048], n_reg from {4, 8, 16}, hyperparameters log-uniform
    (l in [0.3, 3], v in [0.5, 2], noise in [0.01, 0.3]), per-GP seeds; one
    persistent launch fits all of them (assembly, Cholesky, alpha, loss)."""
    from gprat_amd import datagen
    rng = np.random.default_rng(seed)
    probs, flops = [], 0.0
    for g in range(count):
        n = int(np.exp(rng.uniform(np.log(64), np.log(2048))))
        r = int(rng.choice([4, 8, 16]))
        theta = (float(np.exp(rng.uniform(np.log(0.3), np.log(3.0)))),
                 float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))),
                 float(np.exp(rng.uniform(np.log(0.01), np.log(0.3)))))
        u, y = datagen.msd_series(n, 100000 + g)
        probs.append((datagen.padded_input(u, r), y, n, r, theta))
        flops += n ** 3 / 3.0
    out = {"gps": count, "workload": "2048 GPs, n log-uniform [64, 2048], n_reg in {4, 8, 16}"}
    batches = {}
    for name, precision in (("fp64", capi.PRECISION_FP64), ("mixed", capi.PRECISION_MIXED)):
        capi.Batch(ctx, probs[:16], precision=precision).release()  # warm-up
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        b = capi.Batch(ctx, probs, precision=precision)
        wall = time.perf_counter() - t0
        launches, kernel_ms = ctx.kernel_stats("batch_fit")
        ctx.set_profiling(False)
        batches[name] = b
        out[name] = {
            "kernel_ms": kernel_ms,
            "wall_ms_incl_upload": wall * 1e3,
            "gps_per_s": count / (kernel_ms * 1e-3),
            "cholesky_tflops": flops / (kernel_ms * 1e-3) / 1e12,
            "finite_losses": bool(np.all(np.isfinite(b.losses()))),
        }
    # mixed against FP64 on every GP: loss, and mean/variance at 64 points
    l64, l32 = batches["fp64"].losses(), batches["mixed"].losses()
    zt = datagen.padded_input(datagen.msd_series(64, 99)[0], 16)
    em = ev = 0.0
    for g in range(0, count, 16):
        r = probs[g][3]
        z = zt[15 - (r - 1):]
        m64, v64 = batches["fp64"].predict(g, z, 64)
        m32, v32 = batches["mixed"].predict(g, z, 64)
        em = max(em, float(np.max(np.abs(m32 - m64)) / np.max(np.abs(m64))))
        ev = max(ev, float(np.max(np.abs(v32 - v64)) / np.max(np.abs(v64))))
    steps = batches["mixed"].refinement()
    out["mixed"]["speedup_vs_fp64"] = out["fp64"]["kernel_ms"] / out["mixed"]["kernel_ms"]
    out["mixed"]["rel_err_vs_fp64"] = {"mean_every_16th_gp": em, "var_every_16th_gp": ev,
                                       "loss_max": float(np.max(np.abs(l32 - l64) / np.abs(l64)))}
    out["mixed"]["refinement_steps"] = [int(steps.min()), int(steps.max())]
    for b in batches.values():
        b.release()
    return out


def serving_trace(count=192, seed=7):
    """Serving configuration: 70% fit requests (n log-uniform in [256, 32768],
    n_reg = 8, theta = (1, 1, 0.1)), 30% predict requests (n_test uniform in
    [1, 4096]) against a model fitted earlier in the trace (uniform choice);
    per-request seeds."""
    from gprat_amd import datagen
    rng = np.random.default_rng(seed)
    reqs, fits = [], []
    for i in range(count):
        if not fits or rng.random() < 0.7:
            n = int(np.exp(rng.uniform(np.log(256), np.log(32768))))
            u, y = datagen.msd_series(n, 300000 + i)
            reqs.append(("fit", datagen.padded_input(u, 8), y, n))
            fits.append(i)
        else:
            target = int(rng.choice(fits))
            m = int(rng.integers(1, 4097))
            ut, _ = datagen.msd_series(m, 300000 + i)
            reqs.append(("predict", target, datagen.padded_input(ut, 8), m))
    return reqs


def serving_leg(capi, count=192, budget=200 << 30, slots=4):
    """Request loop with continuous batching against the sequential plain
    FP64 path on the same trace.

    * sequential: one request at a time through resident FP64 models (fit =
      assemble + factor + alpha + loss; predict = mean + variance on the
      resident factor). `sequential_refit_req_s` prices the reference's API
      instead, where nothing stays on the device and every predict refits.
    * saturated: the whole trace queued at t = 0 on the server (mixed precision
      for single-GP fits, batched FP64 launches for n <= 2048, 200 GB budget).
    * poisson: open-loop Poisson arrivals at 50% and 80% of the saturated
      rate; latency = completion - arrival.
    """
    import threading
    reqs = serving_trace(count)
    theta = (1.0, 1.0, 0.1)

    # sequential plain FP64, on a context of its own whose cached blocks are
    # freed before the server starts
    seq = capi.Context(0, 1)
    models, times = {}, []
    t_start = time.perf_counter()
    for i, r in enumerate(reqs):
        s = time.perf_counter()
        if r[0] == "fit":
            m = capi.Model(seq, r[1], r[2], r[3], 8, theta, capi.PRECISION_FP64)
            m.loss()
            models[i] = m
        else:
            models[r[1]].predict(r[2], r[3], variance=True)
        times.append(time.perf_counter() - s)
    seq_total = time.perf_counter() - t_start
    refit_total = seq_total + sum(times[r[1]] for r in reqs if r[0] == "predict")
    for m in models.values():
        m.release()
    models.clear()
    seq.close()

    def run(srv, arrivals):
        tickets, submit_t = [None] * len(reqs), [0.0] * len(reqs)
        t0 = time.perf_counter()

        def submitter():
            for i, r in enumerate(reqs):
                if arrivals is not None:
                    d = t0 + arrivals[i] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                submit_t[i] = time.perf_counter()
                if r[0] == "fit":
                    tickets[i] = srv.submit_fit(r[1], r[2], r[3], 8, theta)
                else:
                    tickets[i] = srv.submit_predict(tickets[r[1]], r[2], r[3])

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        lat, bad = [], 0
        for i in range(len(reqs)):
            res = srv.wait(tickets[i])
            # submission to completion, on the server's own clock
            lat.append((res["queue_ms"] + res["service_ms"]) * 1e-3)
            bad += res["status"] != 0
        total = time.perf_counter() - t0
        for i, r in enumerate(reqs):
            if r[0] == "fit":
                srv.release_model(tickets[i])
        return total, lat, bad

    with capi.Server(0, slots=slots, hbm_budget=budget, precision=capi.PRECISION_MIXED) as srv:
        run(srv, None)  # warm-up: device pools and code paths
        sat_total, sat_lat, sat_bad = run(srv, None)
        poisson, po_bad = [], 0
        for frac in (0.5, 0.8):
            rate = frac * len(reqs) / sat_total
            gaps = np.random.default_rng(11).exponential(1.0 / rate, len(reqs))
            po_total, po_lat, bad = run(srv, np.cumsum(gaps) - gaps[0])
            po_bad += bad
            poisson.append({"offered_req_s": rate, "fraction_of_saturated": frac, "achieved_req_s": len(reqs) / po_total,
                            "p50_ms": percentile(po_lat, 50) * 1e3, "p99_ms": percentile(po_lat, 99) * 1e3})
        stats = srv.stats()
    n_fit = sum(r[0] == "fit" for r in reqs)
    # dense work of the trace: n^3/3 per fit, n^2 m per predict (variance
    # solve); the request rate if all of it ran at the matrix-core peaks
    work = sum(r[3] ** 3 / 3.0 for r in reqs if r[0] == "fit") + sum(reqs[r[1]][3] ** 2 * r[3] for r in reqs if r[0] == "predict")
    return {
        "requests": len(reqs),
        "fits": n_fit,
        "predicts": len(reqs) - n_fit,
        "slots": slots,
        "hbm_budget_gb": budget / 2 ** 30,
        "sequential_fp64_req_s": len(reqs) / seq_total,
        "sequential_refit_req_s": len(reqs) / refit_total,
        "saturated_req_s": len(reqs) / sat_total,
        "speedup_vs_sequential": seq_total / sat_total,
        "speedup_vs_sequential_refit": refit_total / sat_total,
        "saturated_p50_ms": percentile(sat_lat, 50) * 1e3,
        "saturated_p99_ms": percentile(sat_lat, 99) * 1e3,
        "poisson": poisson,
        "sequential_p50_ms": percentile(times, 50) * 1e3,
        "sequential_p99_ms": percentile(times, 99) * 1e3,
        "failed": sat_bad + po_bad,
        "trace_flops": work,
        "req_s_if_all_work_at_fp64_peak": len(reqs) / (work / (FP64_MFMA_PEAK_TFLOPS * 1e12)),
        "req_s_if_all_work_at_fp32_peak": len(reqs) / (work / (FP32_MFMA_PEAK_TFLOPS * 1e12)),
        "server_stats": {k: stats[k] for k in ("batched_fits", "batch_launches", "evictions", "refits")},
        "workload": "70% fits n log-uniform [256, 32768] (n_reg 8), 30% predicts m uniform [1, 4096] on earlier models",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--chol-n", type=int, default=32768, help="size of the Cholesky-rate leg (0 skips it)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched many-GP leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop leg")
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
            ("syrk_trailing", "panel", "panel_rows", "cholesky", "assemble", "solve", "trsm", "mean")}
    ctx.set_profiling(False)
    gpu_mean, gpu_var = dmean.numpy(), dvar.numpy()

    per_launch = trailing_update_flops(((n + 127) // 128) * 128)
    launches, syrk_ms = fams["syrk_trailing"]
    expect = len(per_launch) * args.steps
    syrk_flops = sum(per_launch) * args.steps
    achieved = syrk_flops / (syrk_ms * 1e-3) / 1e12 if syrk_ms > 0 else None
    roofline = {
        "kernel": "k_gemm_f64<true,false,5> (Cholesky trailing update from the k-major panel copy, FP64 MFMA)",
        "bound": "mfma",
        "achieved": achieved,
        "peak": FP64_MFMA_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": (achieved / FP64_MFMA_PEAK_TFLOPS) if achieved else None,
        "traffic": measured_traffic(),
        "launches": launches,
        "launches_expected": expect,
        "avg_launch_ms": syrk_ms / launches if launches else None,
        "flops_per_launch_avg": syrk_flops / launches if launches else None,
    }
    chol_ms = fams["cholesky"][1] / max(fams["cholesky"][0], 1)

    calib = ctx.calibrate()
    roofline["peak_measured"] = calib["fp64_tflops"]
    roofline["frac_of_measured"] = (achieved / calib["fp64_tflops"]) if achieved else None

    chol = None
    if args.chol_n > 0:
        chol = cholesky_rate(capi, ctx, args.chol_n)
    mixed = None if args.no_mixed else mixed_leg(capi, ctx)
    mixed64k = None if args.no_mixed else mixed_leg(capi, ctx, n=65536, repeats=1)
    batched = None if args.no_batched else batched_leg(capi, ctx)
    for a in (dx, dy, dz, dmean, dvar):
        a.free()
    ctx.close()  # the serving leg brings its own contexts and HBM budget
    serving = None if args.no_serving else serving_leg(capi)

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
        "mixed_precision_n64k": mixed64k,
        "batched_2048_gps": batched,
        "serving": serving,
        "phase_ms_per_step": {k: v[1] / args.steps for k, v in fams.items()},
        "device_calibration": calib,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        base, err = cpu_baseline(x, y, z, gpu_mean, gpu_var)
        result["cpu_baseline"] = base
        result["rel_err_vs_oracle"] = err
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
