"""The reference's benchmark harness (examples/gprat_cpp/src/execute.cpp)
on this backend: for each host core count (2, 4, ... N_CORES, passed to the
runtime as --hpx:threads) and each training size (START, START*STEP, ...
END), LOOP times: load the data directory through GP_data, build a GP on the
GPU target, start the runtime, time cholesky(), optimize(), and the three
predictions, stop the runtime, and append one row to output.csv with the
reference's columns. Unlike the reference's GPU target, optimize runs on
the GPU here (the reference prints -1).

The data directory holds training_input.txt, training_output.txt and
test_input.txt (the reference's data/data_1024, or one written by
`python -m gprat_amd.datagen DIR N_TRAIN N_TEST`).

Usage: python -m gprat_amd.examples.execute DATA_DIR [--start 512] [--end 1024]
       [--step 2] [--loop 2] [--opt-iter 1] [--n-test 1024] [--cores 4]
       [--n-tiles 16] [--n-reg 8] [--output output.csv]
"""
import argparse
import os
import time

import gprat_amd as gprat

HEADER = ("Target,Cores,N_train,N_test,N_tiles,N_regressor,Opt_iter,Total_time,Init_time,Cholesky_time,Opt_time,"
          "Pred_Uncer_time,Pred_Full_time,Pred_time,N_loop\n")


def run(data_dir, start=512, end=1024, step=2, loop=2, opt_iter=1, n_test=1024, n_cores=4, n_tiles=16, n_reg=8,
        output="output.csv"):
    train_path = os.path.join(data_dir, "training_input.txt")
    out_path = os.path.join(data_dir, "training_output.txt")
    test_path = os.path.join(data_dir, "test_input.txt")
    rows = []
    core = 2
    while core <= n_cores:
        n_train = start
        while n_train <= end:
            for l in range(loop):
                tile_size = gprat.compute_train_tile_size(n_train, n_tiles)
                m_tiles, m_tile_size = gprat.compute_test_tiles(n_test, n_tiles, tile_size)
                hpar = gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, opt_iter)
                training_input = gprat.GP_data(train_path, n_train, n_reg)
                training_output = gprat.GP_data(out_path, n_train, n_reg)
                test_input = gprat.GP_data(test_path, n_test, n_reg)
                t_total = time.perf_counter()
                t0 = time.perf_counter()
                gp = gprat.GP(training_input.data, training_output.data, n_tiles, tile_size, n_reg, [1.0, 1.0, 0.1],
                              [True, True, True], 0, 2)
                init_time = time.perf_counter() - t0
                gprat.start_hpx([], core)
                t0 = time.perf_counter()
                gp.cholesky()
                cholesky_time = time.perf_counter() - t0
                t0 = time.perf_counter()
                gp.optimize(hpar)
                opt_time = time.perf_counter() - t0
                t0 = time.perf_counter()
                gp.predict_with_uncertainty(test_input.data, m_tiles, m_tile_size)
                pred_uncer_time = time.perf_counter() - t0
                t0 = time.perf_counter()
                gp.predict_with_full_cov(test_input.data, m_tiles, m_tile_size)
                pred_full_time = time.perf_counter() - t0
                t0 = time.perf_counter()
                gp.predict(test_input.data, m_tiles, m_tile_size)
                pred_time = time.perf_counter() - t0
                gprat.stop_hpx()
                total_time = time.perf_counter() - t_total
                rows.append(("gpu", core, n_train, n_test, n_tiles, n_reg, opt_iter, total_time, init_time, cholesky_time,
                             opt_time, pred_uncer_time, pred_full_time, pred_time, l))
            n_train *= step
        core *= 2
    new_file = not os.path.exists(output) or os.path.getsize(output) == 0
    with open(output, "a") as f:
        if new_file:
            f.write(HEADER)
        for r in rows:
            f.write(",".join(str(v) for v in r) + "\n")
    return rows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("data_dir")
    ap.add_argument("--start", type=int, default=512)
    ap.add_argument("--end", type=int, default=1024)
    ap.add_argument("--step", type=int, default=2)
    ap.add_argument("--loop", type=int, default=2)
    ap.add_argument("--opt-iter", type=int, default=1)
    ap.add_argument("--n-test", type=int, default=1024)
    ap.add_argument("--cores", type=int, default=4)
    ap.add_argument("--n-tiles", type=int, default=16)
    ap.add_argument("--n-reg", type=int, default=8)
    ap.add_argument("--output", default="output.csv")
    a = ap.parse_args()
    for r in run(a.data_dir, a.start, a.end, a.step, a.loop, a.opt_iter, a.n_test, a.cores, a.n_tiles, a.n_reg, a.output):
        print(",".join(str(v) for v in r))


if __name__ == "__main__":
    main()
