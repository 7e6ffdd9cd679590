"""Autotuner of the batched fit: per size class, which path (persistent
kernel, one workgroup per GP; or level-synchronous launches over all large
GPs, one workgroup per 64x64 block or per 128x128 block) and which precision
(FP64, or mixed when it is faster) the batched fit uses.

For each class of padded sizes (n_pad in (lo, hi], steps of 128 up to 2048)
it fits a batch of GPs of that size (n = hi - 7, n_reg 8, theta (1, 1, 0.1),
seeded data) five ways (FP64: persistent, level, level128; mixed:
persistent, level; the 128-wide steps are FP64 only), each after two
warm-up fits of the same batch (the second captures the level path's launch
sequence, which the timed fit replays), and times the whole batch on the
device with HIP events. The table it writes, one line per class,

    n_pad_max path mixed_faster fp64_persistent_ms fp64_level_ms fp64_level128_ms mixed_persistent_ms mixed_level_ms

is what the library reads from gprat_amd/tuning/batch_<arch>.txt.

Usage: python -m gprat_amd.autotune [OUTPUT] [--gps G]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np


def _batch_ms(capi, ctx, probs, precision, path):
    os.environ["GPRAT_BATCH_PATH"] = path
    os.environ["GPRAT_BATCH_MIXED"] = "requested"
    try:
        for _ in range(2):
            capi.Batch(ctx, probs, precision=precision).release()
        ctx.set_profiling(True)
        b = capi.Batch(ctx, probs, precision=precision)
        _, ms = ctx.kernel_stats("batch")
        ctx.set_profiling(False)
        b.release()
        return ms
    finally:
        os.environ.pop("GPRAT_BATCH_PATH", None)
        os.environ.pop("GPRAT_BATCH_MIXED", None)


def tune(gps_per_class: int = 256, step: int = 128, n_max: int = 2048):
    from gprat_amd import capi, datagen
    ctx = capi.Context(0, 2)
    rows = []
    try:
        for hi in range(step, n_max + 1, step):
            n = hi - 7
            probs = []
            for g in range(gps_per_class):
                u, y = datagen.msd_series(n, 500000 + g)
                probs.append((datagen.padded_input(u, 8), y, n, 8, (1.0, 1.0, 0.1)))
            t = {}
            for path in ("persistent", "level", "level128"):
                t[("fp64", path)] = _batch_ms(capi, ctx, probs, capi.PRECISION_FP64, path)
            for path in ("persistent", "level"):
                t[("mixed", path)] = _batch_ms(capi, ctx, probs, capi.PRECISION_MIXED, path)
            rows.append([hi, None, None, t[("fp64", "persistent")], t[("fp64", "level")], t[("fp64", "level128")],
                         t[("mixed", "persistent")], t[("mixed", "level")]])
            print(f"n_pad <= {hi}: FP64 persistent {t[('fp64', 'persistent')]:.2f} ms, level {t[('fp64', 'level')]:.2f} ms, "
                  f"level128 {t[('fp64', 'level128')]:.2f} ms; mixed persistent {t[('mixed', 'persistent')]:.2f} ms, "
                  f"level {t[('mixed', 'level')]:.2f} ms", flush=True)
    finally:
        ctx.close()
    return choose(rows)


def choose(rows):
    """Per class: persistent or one level variant, and whether mixed beats
    FP64. The level variant is chosen once for all classes (the one whose
    classes, each taking the better of it and the persistent kernel, add up
    to less time): GPs of two level variants would run as two chains of
    launches one after the other."""
    def total(col):
        return sum(min(r[3], r[col]) for r in rows)
    col, name = (4, "level") if total(4) <= total(5) else (5, "level128")
    for r in rows:
        r[1] = name if r[col] < r[3] else "persistent"
        r[2] = int(min(r[6], r[7]) < min(r[3], r[col]))
    return [tuple(r) for r in rows]


def write_table(rows, path, gps_per_class, device_name):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("# batched-fit tuning table (python -m gprat_amd.autotune)\n")
        f.write(f"# device: {device_name}; {gps_per_class} GPs per class (n = n_pad_max - 7, n_reg 8), "
                f"whole-batch time from HIP events, {time.strftime('%Y-%m-%d')}\n")
        f.write("# n_pad_max path mixed_faster fp64_persistent_ms fp64_level_ms fp64_level128_ms mixed_persistent_ms "
                "mixed_level_ms\n")
        for r in rows:
            f.write(f"{r[0]} {r[1]} {r[2]} {r[3]:.3f} {r[4]:.3f} {r[5]:.3f} {r[6]:.3f} {r[7]:.3f}\n")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("output", nargs="?", default=None)
    ap.add_argument("--gps", type=int, default=256)
    args = ap.parse_args()
    from gprat_amd import capi
    name = capi.device_describe(0) if hasattr(capi, "device_describe") else "device 0"
    rows = tune(args.gps)
    out = args.output or os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "batch_gfx950.txt")
    write_table(rows, out, args.gps, name)
    print("wrote", out)


if __name__ == "__main__":
    sys.exit(main())
