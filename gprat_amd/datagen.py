"""Seeded synthetic data for the GP workloads (mass-spring-damper series).

The reference ships fixed data files and an unseeded simulator; the
benchmark configurations instead use this generator so C++, Python and
every run agree on the inputs:

* u_t ~ U[-1, 1] i.i.d. from SplitMix64 (top 53 bits -> [0, 1)).
* Mass-spring-damper, m = 1, k = 1, c = 0.5, dt = 0.05, semi-implicit Euler:
      v <- v + dt * (u_t - k x - c v) / m ;  x <- x + dt * v
  y_t = x (after the update) + N(0, 0.01^2) noise (Box-Muller on the same
  SplitMix64 stream, drawn after all inputs).
* Features: the input vector handed to the GP is the zero-front-padded
  series [0]*(n_reg-1) + u, so sample t's window ends at u_t; y is unpadded.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """`count` SplitMix64 outputs for state `seed` (state advanced first)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int) -> np.ndarray:
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def msd_series(n: int, seed: int, noise_std: float = 0.01):
    """Return (u, y) of length n for one seed."""
    r = uniform01(seed, n + 2 * n)
    u = 2.0 * r[:n] - 1.0
    u1 = r[n:2 * n]
    u2 = r[2 * n:3 * n]
    noise = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    m, k, c, dt = 1.0, 1.0, 0.5, 0.05
    x = v = 0.0
    y = np.empty(n)
    for t in range(n):
        v = v + dt * (u[t] - k * x - c * v) / m
        x = x + dt * v
        y[t] = x
    return u, y + noise_std * noise


def padded_input(u: np.ndarray, n_reg: int) -> np.ndarray:
    return np.concatenate([np.zeros(n_reg - 1), np.asarray(u, dtype=np.float64)])


def gp_problem(n_train: int, n_test: int, n_reg: int, seed_train: int = 0, seed_test: int = 1):
    """Inputs of one GP problem: (train_in, train_out, test_in) as float64
    arrays, train_in/test_in padded with n_reg - 1 leading zeros."""
    u, y = msd_series(n_train, seed_train)
    ut, _ = msd_series(max(n_test, 1), seed_test)
    return padded_input(u, n_reg), y, padded_input(ut[:n_test], n_reg)


def write_dataset(directory: str, n_train: int, n_test: int, seed_train: int = 0, seed_test: int = 1) -> None:
    """Write a seeded dataset in the layout and text format of the
    reference's data directories (data/data_1024: training_input.txt,
    training_output.txt, test_input.txt, test_output.txt, one value per line,
    read by GP_data / utils::load_data with "%lf", core/src/utils_c.cpp:54-81).
    The series are unpadded on disk; GP_data adds the n_reg - 1 leading
    zeros when it loads them."""
    import os
    os.makedirs(directory, exist_ok=True)
    u, y = msd_series(n_train, seed_train)
    ut, yt = msd_series(n_test, seed_test)
    for name, values in (("training_input", u), ("training_output", y), ("test_input", ut), ("test_output", yt)):
        with open(os.path.join(directory, name + ".txt"), "w") as f:
            f.writelines(f"{v!r}\n" for v in np.asarray(values, dtype=np.float64).tolist())


if __name__ == "__main__":
    import sys
    if len(sys.argv) not in (4, 6):
        sys.exit("usage: python -m gprat_amd.datagen DIRECTORY N_TRAIN N_TEST [SEED_TRAIN SEED_TEST]")
    seeds = (int(sys.argv[4]), int(sys.argv[5])) if len(sys.argv) == 6 else (0, 1)
    write_dataset(sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), *seeds)
