"""Writes oracle results at the benchmark configurations that are too slow to
recompute on every test run (minutes of host BLAS), as small .npz fixtures
next to this script. Every value comes from the CPU oracle (oracle/, the
restatement of the reference's cpu:: path) on the seeded generator's data
(gprat_amd.datagen), so the GPU tests compare against the oracle without
running it.

Fixtures (float64 arrays; load with numpy.load, allow_pickle=False):

* oracle_n32768.npz — n_train = 32768 (32 tiles of 1024), n_reg = 16,
  theta = (1, 1, 0.05), the first 2048 test points (2 tiles of 1024):
  mean, var (predict_with_uncertainty), loss (calculate_loss), grad (the
  softplus-space gradient at theta, recovered from the Adam first moment
  after one optimize_step: m_1 = (1 - beta1) g). The test points are a
  prefix of any longer test series of the generator, and a test point's
  mean and variance do not depend on the other test points, so these also
  check the first 2048 outputs of an 8192-point request.
* oracle_n16384.npz — n_train = 16384 (32 tiles of 512), n_reg = 8,
  theta = (1, 1, 0.1): grad (as above), loss, and a 10-iteration optimize
  with AdamParams(0.1, 0.9, 0.999, 1e-8, 10): opt_losses, opt_theta.

Usage: python tests/golden/make_oracle_vectors.py [n32768|n16384 ...]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle", "_build"))

import gprat_oracle  # noqa: E402
from gprat_amd import datagen  # noqa: E402

BETA1 = 0.9


def _grad(gp_factory):
    gp = gp_factory()
    loss = gp.optimize_step(gprat_oracle.AdamParams(0.1, BETA1, 0.999, 1e-8, 1), 0)
    return loss, np.array(gp.adam_moments[0]) / (1.0 - BETA1)


def n32768():
    T, b, r, m = 32, 1024, 16, 2048
    theta = [1.0, 1.0, 0.05]
    x, y, z = datagen.gp_problem(T * b, m, r)
    make = lambda: gprat_oracle.GP(list(x), list(y), T, b, r, theta, [True, True, True])  # noqa: E731
    t0 = time.time()
    gp = make()
    mt, mb = gprat_oracle.compute_test_tiles(m, T, b)
    mean, var = gp.predict_with_uncertainty(list(z), mt, mb)
    loss = gp.calculate_loss() if hasattr(gp, "calculate_loss") else gp.compute_loss()
    t1 = time.time()
    step_loss, grad = _grad(make)
    t2 = time.time()
    print(f"n32768: predict+loss {t1 - t0:.1f} s, gradient {t2 - t1:.1f} s; loss {loss!r} (step {step_loss!r}) grad {grad}",
          flush=True)
    np.savez(os.path.join(HERE, "oracle_n32768.npz"), mean=np.asarray(mean), var=np.asarray(var), loss=np.float64(loss),
             grad=grad, theta=np.asarray(theta), shape=np.array([T, b, r, m]))


def n16384():
    T, b, r = 32, 512, 8
    theta = [1.0, 1.0, 0.1]
    x, y, _ = datagen.gp_problem(T * b, 1, r)
    make = lambda: gprat_oracle.GP(list(x), list(y), T, b, r, theta, [True, True, True])  # noqa: E731
    t0 = time.time()
    step_loss, grad = _grad(make)
    t1 = time.time()
    gp = make()
    losses = gp.optimize(gprat_oracle.AdamParams(0.1, BETA1, 0.999, 1e-8, 10))
    t2 = time.time()
    print(f"n16384: gradient {t1 - t0:.1f} s, 10-iteration optimize {t2 - t1:.1f} s; grad {grad} losses {losses}",
          flush=True)
    np.savez(os.path.join(HERE, "oracle_n16384.npz"), loss=np.float64(step_loss), grad=grad,
             opt_losses=np.asarray(losses), opt_theta=np.asarray(gp.hyperparameters), theta=np.asarray(theta),
             shape=np.array([T, b, r]))


if __name__ == "__main__":
    which = sys.argv[1:] or ["n16384", "n32768"]
    for w in which:
        globals()[w]()
