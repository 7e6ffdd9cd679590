"""The benchmark configurations' API sequences on the GPU, against the oracle.

* n = 1024 in 4 tiles of 256, n_reg = 8, 1024 test points: cholesky(),
  predict_with_uncertainty, calculate_loss and a 3-step Adam optimize at
  lr = 0.1 through the gprat facade (GPU target), each against the oracle
  run live (the reference's own sequence, test/src/output_correctness.cpp:57-112,
  at this size).
* n = 16384 in 32 tiles of 512, n_reg = 8: the loss and softplus-space
  gradient at theta = (1, 1, 0.1) (one optimize_step of the reference) and a
  10-iteration optimize, against tests/golden/oracle_n16384.npz (oracle
  output written by tests/golden/make_oracle_vectors.py; the oracle needs
  minutes of host BLAS here).

Bars: cholesky tiles, mean, variance and loss 1e-10 (normwise / relative);
the gradient 1e-8 relative; Adam losses after 3 steps 1e-10 and hyperparameters 1e-9,
and both after 10 steps at n = 16384 1e-9 (the oracle forms K^-1 by two
dense tiled solves and alpha = K^-1 y by a product, the GPU path by W^T W
and triangular solves: rounding-level differences that the Adam steps
carry along).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, max_rel, norm_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gprat_amd import capi
    c = capi.Context(0, 1)
    yield c
    c.close()


def _theta(gp):
    k = gp.kernel_params
    return [k.lengthscale, k.vertical_lengthscale, k.noise_variance]


def test_c1_sequence(oracle):
    import gprat_amd as gprat
    from gprat_amd import datagen
    T, b, r, m = 4, 256, 8, 1024
    theta = [1.0, 1.0, 0.1]
    x, y, z = datagen.gp_problem(T * b, m, r)
    mt, mb = gprat.compute_test_tiles(m, T, b)
    assert (mt, mb) == oracle.compute_test_tiles(m, T, b) == (4, 256)
    go = oracle.GP(list(x), list(y), T, b, r, theta, [True, True, True])
    gg = gprat.GP(list(x), list(y), T, b, r, theta, [True, True, True], 0, 1)
    tiles_g, tiles_o = gg.cholesky(), go.cholesky()
    worst = max(norm_rel(a, o) for a, o in zip(tiles_g, tiles_o) if len(o))
    assert all(len(a) == len(o) for a, o in zip(tiles_g, tiles_o))
    gm, gv = gg.predict_with_uncertainty(list(z), mt, mb)
    om, ov = go.predict_with_uncertainty(list(z), mt, mb)
    gl, ol = gg.compute_loss(), go.compute_loss()
    lg = gg.optimize(gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, 3))
    lo = go.optimize(oracle.AdamParams(0.1, 0.9, 0.999, 1e-8, 3))
    print(f"n=1024: tiles {worst:.2e} mean {norm_rel(gm, om):.2e} var {norm_rel(gv, ov):.2e} "
          f"loss {abs(gl - ol) / abs(ol):.2e} adam losses {max_rel(lg, lo):.2e} theta {max_rel(_theta(gg), go.hyperparameters):.2e}")
    assert worst < 1e-10
    assert norm_rel(gm, om) < 1e-10 and norm_rel(gv, ov) < 1e-10
    assert abs(gl - ol) <= 1e-10 * abs(ol)
    assert max_rel(lg, lo) < 1e-10
    assert max_rel(_theta(gg), go.hyperparameters) < 1e-9


@pytest.fixture(scope="module")
def n16384():
    return np.load(os.path.join(GOLDEN, "oracle_n16384.npz"))


def test_c2_gradient(ctx, n16384):
    from gprat_amd import datagen
    T, b, r = (int(v) for v in n16384["shape"])
    theta = tuple(float(v) for v in n16384["theta"])
    x, y, _ = datagen.gp_problem(T * b, 1, r)
    gl, gg = ctx.loss_gradient(x, y, theta, T, b, r)
    og, ol = n16384["grad"], float(n16384["loss"])
    rel = np.max(np.abs(np.asarray(gg) - og) / np.abs(og))
    print(f"n=16384 gradient {gg} oracle {og}: rel {rel:.2e}; loss rel {abs(gl - ol) / abs(ol):.2e}")
    assert abs(gl - ol) <= 1e-10 * abs(ol)
    assert rel < 1e-8


def test_c2_optimize_ten_iterations(n16384):
    import gprat_amd as gprat
    from gprat_amd import datagen
    T, b, r = (int(v) for v in n16384["shape"])
    theta = [float(v) for v in n16384["theta"]]
    x, y, _ = datagen.gp_problem(T * b, 1, r)
    gp = gprat.GP(list(x), list(y), T, b, r, theta, [True, True, True], 0, 1)
    losses = gp.optimize(gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, 10))
    el, et = max_rel(losses, n16384["opt_losses"]), max_rel(_theta(gp), n16384["opt_theta"])
    print(f"n=16384 optimize(10): losses {el:.2e} theta {et:.2e}")
    assert len(losses) == 10
    assert el < 1e-9 and et < 1e-9
