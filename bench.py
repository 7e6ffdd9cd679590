#!/usr/bin/env python3
"""Benchmark of the GP fit+predict path on one MI355X.

A step is one fit+predict request at the mixed-precision configuration:
n_train = 32768 (32 tiles of 1024), n_reg = 16, n_test = 8192,
theta = (1, 1, 0.05), seeded mass-spring-damper data. The request assembles
K, factors it with FP64 diagonal blocks and FP32 matrix-core trailing
updates, solves alpha = K^-1 y with FP64 iterative refinement, and returns
the predictive mean and the latent variance, refined in FP64 (the
reference's predict_with_uncertainty). Inputs are placed in HBM before the
timed region and outputs stay there, so `value` excludes PCIe.

Also reported on the same line:
  * p50 / p99 request latency over the timed steps, and the request's error
    against the oracle (mean and variance of the first 2048 test points,
    loss and gradient; tests/golden/oracle_n32768.npz);
  * the same request in plain FP64 and with the variance unrefined (FP32
    solve only), and the speed-up of each mixed variant over FP64;
  * the plain-FP64 scale configuration (n_train = 16384 in 32 tiles of 512,
    n_reg 8, n_test 4096) as a request rate with p50 / p99, and its
    10-iteration Adam optimize;
  * the n = 32768 FP64 Cholesky rate, the n = 65536 mixed-precision request,
    the batched many-GP fit and the serving loop (median of three runs);
  * `roofline` for the request's dominant kernel, from HIP events bracketing
    each of its launches in the timed region;
  * `cpu_baseline`: the CPU oracle (a restatement of the reference's HPX +
    BLAS path, oneMKL sequential per tile on a host pool) on a bounded
    sample of the workload, scaled to a whole request by its flops.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--headline-only]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fit+predict req/s & p50/p99 latency at ≤1e-8 rel err; n=32k Cholesky TFLOP/s"
FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

# the request of the headline (mixed-precision configuration, n = 32768)
HN, HT, HR, HM = 32768, 32, 16, 8192
HTHETA = (1.0, 1.0, 0.05)
# the plain-FP64 scale configuration
N_TILES, TILE, N_REG, N_TEST = 32, 512, 8, 4096
THETA = (1.0, 1.0, 0.1)


def outer_block(n_pad: int) -> int:
    env = os.environ.get("GPRAT_CHOLESKY_BLOCK")
    if env and int(env) >= 128 and int(env) % 128 == 0:
        return int(env)
    return 256 if n_pad <= 4096 else (512 if n_pad < 12288 else 768)


def trailing_update_flops(n_pad: int, extra: int = 0):
    """Algorithmic flops of each trailing-update launch, in launch order:
    2 * K * (number of elements updated). Step k's launch applies panel k
    (K = its width) to the rows below the next diagonal block: the next
    panel's columns whole, the columns after it lower (the next diagonal
    block itself is updated on the panel stream); with `extra` rows below
    the matrix (the one-call prediction's test rows), also every column of
    those rows right of the panel."""
    bo = outer_block(n_pad)
    out = []
    p0 = 0
    while p0 < n_pad:
        pe = min(n_pad, p0 + bo)
        if pe >= n_pad:
            break
        q1 = min(n_pad, pe + bo)
        K = pe - p0
        S = n_pad - q1
        f = 2.0 * K * (S * (q1 - pe) + S * (S + 1) / 2 + extra * (n_pad - pe))
        if q1 < n_pad or extra > 0:
            out.append(f)
        p0 = pe
    return out


def residual_flops(n_pad: int, m_pad: int) -> float:
    """Algorithmic flops of the variance refinement's FP64 product Z = K_sl X:
    every 128-row tile t takes the strictly lower part left of its diagonal
    block, 2 * (128 t) * 128 * m_pad flops."""
    T = n_pad // 128
    return 2.0 * m_pad * 128 * 128 * T * (T - 1) / 2.0


def mixed_request_flops(n_pad: int, m_pad: int) -> dict:
    """Algorithmic flops of one mixed-precision request (resident fit, then a
    prediction with the refined variance) by the precision they run in. The
    Cholesky's panel of width w at p0 is factored and solved for the rows
    below it in FP64 (w^3/3 + (n - p0 - w) w^2); the rest of n^3/3, the
    trailing update, runs on FP32 matrix cores. The variance's forward and
    backward solves are n^2 m each on FP32 matrix cores, its residual
    product (the strictly lower 128-row blocks of K times X) FP64. The
    O(n^2) terms (alpha, its refinement, the mean) are left out."""
    bo = outer_block(n_pad)
    panels = 0.0
    p0 = 0
    while p0 < n_pad:
        w = min(bo, n_pad - p0)
        panels += w ** 3 / 3.0 + (n_pad - p0 - w) * float(w) * w
        p0 += w
    chol = n_pad ** 3 / 3.0
    return {
        "fp64": panels + residual_flops(n_pad, m_pad),
        "fp32": (chol - panels) + 2.0 * n_pad ** 2 * m_pad,
        "fp64_cholesky_panels": panels,
        "fp32_cholesky_trailing": chol - panels,
        "fp32_variance_solves": 2.0 * n_pad ** 2 * m_pad,
        "fp64_variance_residual": residual_flops(n_pad, m_pad),
        "panel_width": bo,
    }


def variance_update_flops(n_pad: int, m_pad: int, nbk: int, directions: int):
    """Algorithmic flops of the variance solves' update products over one
    request: step J of the forward solve updates the n_pad - j1 rows below
    its block, step J of the backward solve the j0 rows above it, each a
    (rows x m_pad) += (rows x nb) (nb x m_pad) product; returns (total flops,
    launches)."""
    total, launches = 0.0, 0
    j0 = 0
    while j0 < n_pad:
        nb = min(nbk, n_pad - j0)
        j1 = j0 + nb
        if j1 < n_pad:
            total += 2.0 * (n_pad - j1) * nb * m_pad
            launches += 1
        if directions == 2 and j0 > 0:
            total += 2.0 * j0 * nb * m_pad
            launches += 1
        j0 = j1
    return total, launches


def measured_traffic(name):
    """HBM bytes per launch of a kernel from a committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a 1 GiB copy); None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def percentile(values, q):
    return float(np.percentile(np.asarray(values), q))


def host_description():
    model = ""
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(repeats=5):
    """The oracle on a bounded sample of the headline workload: fit +
    predict_with_uncertainty at n = 16384 (16 tiles of 1024, n_reg 16, noise
    0.05) on 1024 test points, one warm-up and `repeats` timed runs, median.
    A whole request (n = 32768, m = 8192) does F = n^3/3 + n^2 m flops of
    tile BLAS against the sample's; `value` scales the sample's median time
    by that ratio."""
    sys.path.insert(0, os.path.join(ROOT, "oracle", "_build"))
    import gprat_oracle
    from gprat_amd import datagen
    nproc, model = host_description()
    threads = min(nproc, 16)
    gprat_oracle.set_threads(threads)
    T, b, r, m = 16, 1024, HR, 1024
    x, y, z = datagen.gp_problem(T * b, m, r)
    mt, mb = gprat_oracle.compute_test_tiles(m, T, b)
    times = []
    for rep in range(repeats + 1):
        gp = gprat_oracle.GP(list(x), list(y), T, b, r, list(HTHETA), [True, True, True])
        t0 = time.perf_counter()
        gp.predict_with_uncertainty(list(z), mt, mb)
        dt = time.perf_counter() - t0
        if rep > 0:
            times.append(dt)
    med = float(np.median(times))
    ns = T * b
    scale = (HN ** 3 / 3.0 + HN ** 2 * HM) / (ns ** 3 / 3.0 + ns ** 2 * m)
    return {
        "value": 1.0 / (med * scale),
        "unit": "req/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle fit + predict_with_uncertainty at n={ns} (16 tiles of 1024, n_reg {r}), m={m}: "
                  f"1 warm-up + {repeats} runs, median {med:.2f} s ({', '.join(f'{t:.2f}' for t in times)}); "
                  f"scaled by the flop ratio {scale:.2f} to one n={HN}, m={HM} request; "
                  f"{gprat_oracle.blas_name()} sequential per tile, {threads} pool threads (cap 16)",
        "sample_req_s": 1.0 / med,
        "nproc": nproc,
        "cpu_model": model,
    }


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


def request_leg(capi, ctx, n, T, r, m, theta, precision, steps, warmup=1, refine_variance=True, families=(), oneshot=False):
    """Times `steps` fit+predict requests (inputs resident in HBM): latency
    per request and, for `families`, the kernel-family stats of the timed
    requests. A request is a resident model's fit and predict, or with
    `oneshot` the one-call prediction (gprat_predict_with_uncertainty, FP64,
    the reference's own call). Returns the timing and a model of the same
    problem (fitted after the timed region; released by the caller)."""
    from gprat_amd import datagen
    x, y, z = datagen.gp_problem(n, m, r)
    dx, dy, dz = ctx.device_array(x), ctx.device_array(y), ctx.device_array(z)
    dmean, dvar = ctx.device_empty(m), ctx.device_empty(m)

    def step(keep=False):
        if oneshot and not keep:
            ctx.predict(dx, dy, dz, theta, T, n // T, 1, m, r, variance=True, out=(dmean, dvar))
            return None
        mdl = capi.Model(ctx, dx, dy, n, r, theta, precision)
        if oneshot:
            return mdl
        if not refine_variance:
            mdl.set_variance_refinement(False)
        mdl.predict(dz, m, variance=True, out=(dmean, dvar))
        if keep:
            return mdl
        mdl.release()
        return None

    for _ in range(warmup):
        step()
    ctx.synchronize()
    if families:
        ctx.set_profiling(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        step()
        lat.append(time.perf_counter() - s)
    ctx.synchronize()
    total = time.perf_counter() - t0
    fams = {f: ctx.kernel_stats(f) for f in families}
    if families:
        ctx.set_profiling(False)
    model = step(keep=True)
    out = {
        "req_s": steps / total,
        "ms_per_request": total / steps * 1e3,
        "p50_ms": percentile(lat, 50) * 1e3,
        "p99_ms": percentile(lat, 99) * 1e3,
        "steps": steps,
        "workload": f"fit+predict n={n} ({T} tiles of {n // T}), n_reg={r}, m={m}, theta={theta}",
    }
    mean, var = dmean.numpy(), dvar.numpy()
    for a in (dx, dy, dz, dmean, dvar):
        a.free()
    return out, fams, model, (mean, var, x, y, z)


def c2_oneshot_leg(capi, ctx, steps, warmup):
    """The C2 request (n = 16384 in 32 tiles of 512, n_reg 8, m = 4096,
    FP64) through the reference's own call, gprat_predict_with_uncertainty:
    fit and predict in one, the test rows of K* factored below K."""
    n16 = N_TILES * TILE
    fams = ("syrk_trailing", "cholesky", "panel", "panel_rows", "assemble", "solve", "variance_rows", "mean")
    c2, f, mdl, _ = request_leg(capi, ctx, n16, N_TILES, N_REG, N_TEST, THETA, capi.PRECISION_FP64, steps, warmup=warmup,
                                families=fams, oneshot=True)
    mdl.release()
    c2["api"] = "gprat_predict_with_uncertainty (one call: fit + predict, test rows and targets factored with K)"
    m_pad = ((N_TEST + 63) // 64) * 64
    tr = trailing_update_flops(n16, m_pad + 64)  # the test rows, then the targets' 64-row block
    c2["roofline_trailing_update"] = roofline_of(
        "k_gemm_f64<true,false,5> (FP64 trailing update, test rows included)", sum(tr) / len(tr), f["syrk_trailing"],
        FP64_MFMA_PEAK_TFLOPS, "trailing_update_oneshot_n16384_traffic.md")
    fact_ms = f["cholesky"][1] / max(f["cholesky"][0], 1)
    c2["factorisation_with_test_rows_ms"] = fact_ms
    c2["factorisation_with_test_rows_tflops"] = (n16 ** 3 / 3.0 + n16 ** 2 * N_TEST) / (fact_ms * 1e-3) / 1e12
    c2["phase_ms_per_request"] = {k: v[1] / steps for k, v in f.items()}
    return c2


def c2_resident_leg(capi, ctx, steps, warmup):
    """The same request through a resident model (gprat_model_fit, then
    gprat_model_predict), whose factorisation is the plain Cholesky."""
    n16 = N_TILES * TILE
    fams = ("syrk_trailing", "cholesky", "panel", "panel_rows", "assemble", "solve", "trsm", "mean")
    c2r, f, mdl, _ = request_leg(capi, ctx, n16, N_TILES, N_REG, N_TEST, THETA, capi.PRECISION_FP64, steps, warmup=warmup,
                                 families=fams)
    mdl.release()
    chol_ms = f["cholesky"][1] / max(f["cholesky"][0], 1)
    c2r["api"] = "resident model: gprat_model_fit + gprat_model_predict"
    c2r["cholesky_ms"] = chol_ms
    c2r["cholesky_tflops"] = n16 ** 3 / 3.0 / (chol_ms * 1e-3) / 1e12
    tr = trailing_update_flops(n16)
    c2r["roofline_trailing_update"] = roofline_of(
        "k_gemm_f64<true,false,5> (FP64 trailing update)", sum(tr) / len(tr), f["syrk_trailing"], FP64_MFMA_PEAK_TFLOPS,
        "trailing_update_n16384_traffic.md")
    c2r["phase_ms_per_request"] = {k: v[1] / steps for k, v in f.items()}
    return c2r


def vs_oracle_n32768(mean, var, loss, grad):
    """Errors of the first 2048 test points' mean and variance, the loss and
    the gradient against the oracle at n = 32768 (committed fixture)."""
    ref = np.load(os.path.join(ROOT, "tests", "golden", "oracle_n32768.npz"))
    k = len(ref["mean"])
    return {
        "mean": float(np.max(np.abs(mean[:k] - ref["mean"])) / np.max(np.abs(ref["mean"]))),
        "var": float(np.max(np.abs(var[:k] - ref["var"])) / np.max(np.abs(ref["var"]))),
        "loss_abs": abs(loss - float(ref["loss"])),
        "loss_rel": abs(loss - float(ref["loss"])) / abs(float(ref["loss"])),
        "grad_rel_max": float(np.max(np.abs(np.asarray(grad) - ref["grad"]) / np.abs(ref["grad"]))),
        "points": int(k),
        "norm": "max|a-b| / max|b| over the first 2048 test points (mean, var); scalar relative (loss, grad)",
    }


def mixed64k_leg(capi, ctx, n=65536, m=8192, r=16):
    """n = 65536 (64 tiles of 1024): one FP64 and one mixed-precision
    (refined) request, after one warm-up of the mixed one; the mixed
    result's error against the FP64 one."""
    theta = HTHETA
    res = {}
    for name, prec in (("mixed", capi.PRECISION_MIXED), ("fp64", capi.PRECISION_FP64)):
        t, _, mdl, (mean, var, *_rest) = request_leg(capi, ctx, n, 64, r, m, theta, prec, steps=1, warmup=1 if name == "mixed" else 0)
        res[name] = (t, mean, var, mdl.loss())
        mdl.release()
    (t32, m32, v32, l32), (t64, m64, v64, l64) = res["mixed"], res["fp64"]
    return {
        "fp64_ms": t64["ms_per_request"],
        "mixed_ms": t32["ms_per_request"],
        "speedup": t64["ms_per_request"] / t32["ms_per_request"],
        "rel_err_vs_fp64": {"mean": float(np.max(np.abs(m32 - m64)) / np.max(np.abs(m64))),
                            "var": float(np.max(np.abs(v32 - v64)) / np.max(np.abs(v64))),
                            "loss_abs": abs(l32 - l64)},
        "workload": t32["workload"],
    }


def optimize_leg(repeats=1):
    """The plain-FP64 scale configuration's 10-iteration Adam optimize
    (lr 0.1) through the gprat facade on the GPU, against the oracle's
    losses and hyperparameters (committed fixture)."""
    import gprat_amd as gprat
    from gprat_amd import datagen
    ref = np.load(os.path.join(ROOT, "tests", "golden", "oracle_n16384.npz"))
    T, b, r = (int(v) for v in ref["shape"])
    x, y, _ = datagen.gp_problem(T * b, 1, r)
    theta = [float(v) for v in ref["theta"]]
    gp = gprat.GP(list(x), list(y), T, b, r, theta, [True, True, True], 0, 1)
    gp.optimize(gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, 1))  # warm-up
    gp = gprat.GP(list(x), list(y), T, b, r, theta, [True, True, True], 0, 1)
    t0 = time.perf_counter()
    losses = gp.optimize(gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, 10))
    dt = time.perf_counter() - t0
    k = gp.kernel_params
    th = np.array([k.lengthscale, k.vertical_lengthscale, k.noise_variance])
    n = T * b
    return {
        "iterations": 10,
        "ms_total": dt * 1e3,
        "ms_per_iteration": dt * 1e2,
        "rel_err_vs_oracle": {"losses": float(np.max(np.abs(np.asarray(losses) - ref["opt_losses"]) / np.abs(ref["opt_losses"]))),
                              "theta": float(np.max(np.abs(th - ref["opt_theta"]) / np.abs(ref["opt_theta"])))},
        "flops_per_iteration_this_algorithm": n ** 3 / 3.0 * 3,
        "flops_per_iteration_reference_algorithm": n ** 3 / 3.0 + 2.0 * n ** 3,
        # whole iterations (host Adam and transfers included) against the
        # FP64 matrix peak: the executed algorithm's flops, and for
        # comparison the reference algorithm's at the same time
        "tflops_this_algorithm": n ** 3 / 3.0 * 3 / (dt * 1e-1) / 1e12,
        "frac_of_fp64_peak_this_algorithm": n ** 3 / 3.0 * 3 / (dt * 1e-1) / 1e12 / FP64_MFMA_PEAK_TFLOPS,
        "tflops_reference_algorithm_equivalent": (n ** 3 / 3.0 + 2.0 * n ** 3) / (dt * 1e-1) / 1e12,
        "workload": f"optimize: n={n} ({T} tiles of {b}), n_reg={r}, 10 Adam iterations at lr 0.1 (K^-1 = W^T W per iteration)",
    }


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
        # the tuning table routes every size class to FP64 (mixed measured
        # slower in each); the mixed leg forces mixed precision to show why
        if name == "mixed":
            os.environ["GPRAT_BATCH_MIXED"] = "requested"
        # two warm-ups: the second captures the level path's launch sequence
        # for this batch composition, which the timed fit then replays
        for _ in range(2):
            capi.Batch(ctx, probs, precision=precision).release()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        b = capi.Batch(ctx, probs, precision=precision)
        wall = time.perf_counter() - t0
        launches, kernel_ms = ctx.kernel_stats("batch")
        ctx.set_profiling(False)
        batches[name] = b
        out[name] = {
            "kernel_ms": kernel_ms,
            "wall_ms_incl_upload": wall * 1e3,
            "gps_per_s": count / (kernel_ms * 1e-3),
            "cholesky_tflops": flops / (kernel_ms * 1e-3) / 1e12,
            "finite_losses": bool(np.all(np.isfinite(b.losses()))),
        }
    os.environ.pop("GPRAT_BATCH_MIXED", None)
    out["tuning_table"] = "gprat_amd/tuning/batch_gfx950.txt (python -m gprat_amd.autotune)"
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


def serving_leg(capi, count=192, budget=200 << 30, slots=4, runs=3):
    """Request loop with continuous batching against the sequential plain
    FP64 path on the same trace; every figure is the median of `runs` runs.

    * sequential: one request at a time through resident FP64 models (fit =
      assemble + factor + alpha + loss; predict = mean + variance on the
      resident factor). `sequential_refit_req_s` prices the reference's API
      instead, where nothing stays on the device and every predict refits.
    * saturated: the whole trace queued at t = 0 on the server (mixed
      precision with FP64 refinement for fits with n > 2048; smaller fits in
      FP64, batched when several wait together and that pays; 200 GB budget).
    * poisson: open-loop Poisson arrivals at 50% and 80% of the saturated
      rate; latency = completion - arrival.
    """
    import threading
    reqs = serving_trace(count)
    theta = (1.0, 1.0, 0.1)

    def sequential():
        # on a context of its own whose cached blocks are freed afterwards
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
        total = time.perf_counter() - t_start
        refit = total + sum(times[r[1]] for r in reqs if r[0] == "predict")
        for m in models.values():
            m.release()
        seq.close()
        return total, refit, times

    def run(srv, arrivals):
        tickets = [None] * len(reqs)
        t0 = time.perf_counter()

        def submitter():
            for i, r in enumerate(reqs):
                if arrivals is not None:
                    d = t0 + arrivals[i] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
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
        if arrivals is not None:
            # completion times on the arrival clock (arrival + latency)
            done = [arrivals[i] + lat[i] for i in range(len(reqs))]
            return total, lat, bad, done
        return total, lat, bad

    seq_runs = [sequential() for _ in range(runs)]
    seq_total = float(np.median([t for t, _, _ in seq_runs]))
    refit_total = float(np.median([f for _, f, _ in seq_runs]))
    seq_times = seq_runs[int(np.argsort([t for t, _, _ in seq_runs])[len(seq_runs) // 2])][2]
    failed = 0
    with capi.Server(0, slots=slots, hbm_budget=budget, precision=capi.PRECISION_MIXED) as srv:
        run(srv, None)  # warm-up: device pools and code paths
        sat = [run(srv, None) for _ in range(runs)]
        failed += sum(b for _, _, b in sat)
        sat_total = float(np.median([t for t, _, _ in sat]))
        sat_lat = sat[int(np.argsort([t for t, _, _ in sat])[len(sat) // 2])][1]
        poisson = []
        for frac in (0.5, 0.8):
            rate = frac * len(reqs) / sat_total
            gaps = np.random.default_rng(11).exponential(1.0 / rate, len(reqs))
            arrivals = np.cumsum(gaps) - gaps[0]
            po = [run(srv, arrivals) for _ in range(runs)]
            failed += sum(b for _, _, b, _ in po)
            # achieved = requests / makespan (first arrival to last completion):
            # the last large fits' own service follows the last arrival, so it
            # stays below the offered rate even when the server keeps up; the
            # tail (makespan - arrival span) and the share of requests done by
            # the last arrival show whether work piled up
            span = float(arrivals[-1])
            poisson.append({"offered_req_s": rate, "fraction_of_saturated": frac,
                            "arrival_span_s": span,
                            "makespan_s": float(np.median([t for t, _, _, _ in po])),
                            "achieved_req_s": float(np.median([len(reqs) / t for t, _, _, _ in po])),
                            "tail_s": float(np.median([t for t, _, _, _ in po])) - span,
                            "done_by_last_arrival": float(np.median([np.mean(np.asarray(d) <= span) for _, _, _, d in po])),
                            "p50_ms": float(np.median([percentile(l, 50) for _, l, _, _ in po])) * 1e3,
                            "p99_ms": float(np.median([percentile(l, 99) for _, l, _, _ in po])) * 1e3,
                            "p99_ms_runs": [percentile(l, 99) * 1e3 for _, l, _, _ in po]})
        stats = srv.stats()
    n_fit = sum(r[0] == "fit" for r in reqs)
    # dense work of the trace: n^3/3 per fit, n^2 m per predict (variance
    # solve); the request rate if all of it ran at the matrix-core peaks
    work = sum(r[3] ** 3 / 3.0 for r in reqs if r[0] == "fit") + sum(reqs[r[1]][3] ** 2 * r[3] for r in reqs if r[0] == "predict")
    fp32_bound = len(reqs) / (work / (FP32_MFMA_PEAK_TFLOPS * 1e12))
    return {
        "requests": len(reqs),
        "fits": n_fit,
        "predicts": len(reqs) - n_fit,
        "slots": slots,
        "runs": runs,
        "hbm_budget_gb": budget / 2 ** 30,
        "sequential_fp64_req_s": len(reqs) / seq_total,
        "sequential_fp64_req_s_runs": [len(reqs) / t for t, _, _ in seq_runs],
        "sequential_refit_req_s": len(reqs) / refit_total,
        "saturated_req_s": len(reqs) / sat_total,
        "saturated_req_s_runs": [len(reqs) / t for t, _, _ in sat],
        "speedup_vs_sequential": seq_total / sat_total,
        "speedup_vs_sequential_refit": refit_total / sat_total,
        "saturated_p50_ms": percentile(sat_lat, 50) * 1e3,
        "saturated_p99_ms": percentile(sat_lat, 99) * 1e3,
        "poisson": poisson,
        "sequential_p50_ms": percentile(seq_times, 50) * 1e3,
        "sequential_p99_ms": percentile(seq_times, 99) * 1e3,
        "failed": failed,
        "trace_flops": work,
        "req_s_if_all_work_at_fp64_peak": len(reqs) / (work / (FP64_MFMA_PEAK_TFLOPS * 1e12)),
        "req_s_if_all_work_at_fp32_peak": fp32_bound,
        "saturated_fraction_of_fp32_peak_bound": (len(reqs) / sat_total) / fp32_bound,
        "server_stats": {k: stats[k] for k in ("batched_fits", "batch_launches", "evictions", "refits",
                                                        "cache_flushes", "borrowed_blocks")},
        "workload": "70% fits n log-uniform [256, 32768] (n_reg 8), 30% predicts m uniform [1, 4096] on earlier models",
    }


def roofline_of(kernel, flops_per_launch, fam, peak, traffic_file):
    launches, ms = fam
    if not launches or ms <= 0:
        return None
    achieved = flops_per_launch * launches / (ms * 1e-3) / 1e12
    return {
        "kernel": kernel,
        "bound": "mfma",
        "achieved": achieved,
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": measured_traffic(traffic_file),
        "launches": launches,
        "avg_launch_ms": ms / launches,
        "flops_per_launch_avg": flops_per_launch,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="only the timed requests of the headline (for profiling)")
    ap.add_argument("--chol-n", type=int, default=32768, help="size of the Cholesky-rate leg (0 skips it)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched many-GP leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--only", choices=("batched", "serving", "c2", "c2-resident", "chol"), default=None,
                    help="run one leg alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this backend runs on one GPU; the GP path has no multi-GPU mode "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from gprat_amd import capi

    if args.only == "serving":
        print(json.dumps({"serving": serving_leg(capi)}), flush=True)
        return
    ctx = capi.Context(0, 1)
    if args.only == "batched":
        out = batched_leg(capi, ctx)
        ctx.close()
        print(json.dumps({"batched_2048_gps": out}), flush=True)
        return
    if args.only == "chol":
        out = cholesky_rate(capi, ctx, args.chol_n)
        ctx.close()
        print(json.dumps({"cholesky": out}), flush=True)
        return
    if args.only in ("c2", "c2-resident"):
        leg = c2_oneshot_leg if args.only == "c2" else c2_resident_leg
        out = leg(capi, ctx, args.steps, args.warmup)
        ctx.close()
        print(json.dumps({args.only: out}), flush=True)
        return
    families = ("syrk_trailing", "residual", "trsm", "trsm_update", "panel", "panel_rows", "cholesky", "assemble", "solve", "mean",
                "refine")
    head, fams, model, (mean, var, *_r) = request_leg(capi, ctx, HN, HT, HR, HM, HTHETA, capi.PRECISION_MIXED, args.steps,
                                                      warmup=args.warmup, families=families)
    n_pad = ((HN + 127) // 128) * 128
    m_pad = ((HM + 63) // 64) * 64
    trailing = trailing_update_flops(n_pad)
    r_fp32 = roofline_of("k_gemm_f32_tn<double,0> (FP32 matrix-core trailing update of the mixed-precision Cholesky)",
                         sum(trailing) / len(trailing), fams["syrk_trailing"], FP32_MFMA_PEAK_TFLOPS,
                         "trailing_update_f32_n32768_traffic.md")
    r_res = roofline_of("k_gemm_f64<true,false,8> (FP64 residual product K_sl X of the variance refinement)",
                        residual_flops(n_pad, m_pad), fams["residual"], FP64_MFMA_PEAK_TFLOPS, "residual_n32768_traffic.md")
    upd_flops, upd_launches = variance_update_flops(n_pad, m_pad, outer_block(n_pad), 2)
    r_upd = roofline_of("k_gemm_f32_tn<float,2> (FP32 matrix-core updates of the variance's forward and backward solves)",
                        upd_flops / upd_launches, fams["trsm_update"], FP32_MFMA_PEAK_TFLOPS, "trsm_update_f32_n32768_traffic.md")
    ranked = sorted([r for r in (r_fp32, r_res, r_upd) if r], key=lambda r: -r["launches"] * r["avg_launch_ms"])
    roofline = ranked[0] if ranked else None
    loss = model.loss()
    _, grad = model.gradient()
    it, last = model.refinement()
    vcorr = model.variance_correction()
    model.release()
    head["rel_err_vs_oracle"] = vs_oracle_n32768(mean, var, loss, grad)
    head["alpha_refinement_steps"] = it
    head["alpha_last_correction"] = last
    head["variance_max_correction"] = vcorr
    head["phase_ms_per_request"] = {k: v[1] / args.steps for k, v in fams.items()}
    head["flops_by_precision"] = mixed_request_flops(n_pad, m_pad)

    result = {
        "metric": METRIC,
        "value": head["req_s"],
        "unit": "req/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_request"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (FP32 matrix-core trailing updates and solves, FP64 refinement)",
        "data": "synthetic (seeded mass-spring-damper series, SplitMix64)",
        "config": {
            "workload": f"fit+predict request, mixed precision with FP64 refinement of alpha and of the variance: "
                        f"n_train={HN} ({HT} tiles of {HN // HT}), n_reg={HR}, n_test={HM}, theta={HTHETA}, "
                        f"inputs resident in HBM",
            "precision": "mixed",
        },
        "p50_ms": head["p50_ms"],
        "p99_ms": head["p99_ms"],
        "headline": head,
        "roofline": roofline,
        "roofline_other": ranked[1:],
    }
    if not args.headline_only:
        f64, _, mdl, (m64, v64, *_r) = request_leg(capi, ctx, HN, HT, HR, HM, HTHETA, capi.PRECISION_FP64, 3, warmup=1)
        f64["rel_err_vs_oracle"] = vs_oracle_n32768(m64, v64, mdl.loss(), mdl.gradient()[1])
        mdl.release()
        unr, _, mdl, (mu, vu, *_r) = request_leg(capi, ctx, HN, HT, HR, HM, HTHETA, capi.PRECISION_MIXED, 3, warmup=1,
                                                 refine_variance=False)
        unr["rel_err_vs_oracle"] = vs_oracle_n32768(mu, vu, mdl.loss(), mdl.gradient()[1])
        mdl.release()
        one, _, mdl, (mo, vo, *_r) = request_leg(capi, ctx, HN, HT, HR, HM, HTHETA, capi.PRECISION_FP64, 3, warmup=1,
                                                 oneshot=True)
        one["api"] = "gprat_predict_with_uncertainty (one call, FP64, test rows and targets factored with K)"
        one["rel_err_vs_oracle"] = vs_oracle_n32768(mo, vo, mdl.loss(), mdl.gradient()[1])
        mdl.release()
        result["fp64_n32k"] = f64
        result["fp64_n32k_one_call"] = one
        result["mixed_unrefined_variance_n32k"] = unr
        result["speedup_mixed_refined_vs_fp64"] = f64["ms_per_request"] / head["ms_per_request"]
        result["speedup_mixed_unrefined_vs_fp64"] = f64["ms_per_request"] / unr["ms_per_request"]
        result["fp64_n16k"] = c2_oneshot_leg(capi, ctx, 10, 2)
        result["fp64_n16k_resident"] = c2_resident_leg(capi, ctx, 10, 2)
        result["optimize_n16k"] = optimize_leg()
        calib = ctx.calibrate()
        result["device_calibration"] = calib
        if roofline:
            peak_meas = calib["fp64_tflops"] if "f64" in roofline["kernel"] else calib.get("fp32_tflops")
            if peak_meas:
                roofline["peak_measured"] = peak_meas
                roofline["frac_of_measured"] = roofline["achieved"] / peak_meas
        result["cholesky_n32k"] = cholesky_rate(capi, ctx, args.chol_n) if args.chol_n > 0 else None
        result["mixed_precision_n64k"] = mixed64k_leg(capi, ctx)
        result["batched_2048_gps"] = None if args.no_batched else batched_leg(capi, ctx)
    ctx.close()  # the serving leg brings its own contexts and HBM budget
    if not args.headline_only:
        result["serving"] = None if args.no_serving else serving_leg(capi)
    result["cpu_baseline"] = None if (args.no_cpu_baseline or args.headline_only) else cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
