"""The reference's own integration test, run on the GPU through the drop-in
Python API: test/src/output_correctness.cpp:57-112 (config) and :191-248
(checks), tolerance WithinRel(eps * 1e6) = 2.22e-10. Unlike the reference's
GPU test (:251-309), the loss/optimizer leg is included: it runs on the GPU
here."""
import json
import os

import numpy as np
import pytest

from conftest import within_rel, max_rel

TOL = np.finfo(np.float64).eps * 1e6

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def results(golden_dir):
    import gprat_amd as gprat
    n_train = n_test = 128
    n_tiles, n_reg = 4, 8
    ts = gprat.compute_train_tile_size(n_train, n_tiles)
    mt = gprat.compute_test_tiles(n_test, n_tiles, ts)
    xi = gprat.GP_data(os.path.join(golden_dir, "training_input.txt"), n_train, n_reg)
    yo = gprat.GP_data(os.path.join(golden_dir, "training_output.txt"), n_train, n_reg)
    zi = gprat.GP_data(os.path.join(golden_dir, "test_input.txt"), n_test, n_reg)
    gp = gprat.GP(xi.data, yo.data, n_tiles, ts, n_reg, [1.0, 1.0, 0.1], [True, True, True], 0, 1)
    out = {
        "choleksy": gp.cholesky(),
        "sum": gp.predict_with_uncertainty(zi.data, *mt),
        "full": gp.predict_with_full_cov(zi.data, *mt),
        "pred": gp.predict(zi.data, *mt),
        "losses": gp.optimize(gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, 3)),
    }
    out["theta"] = [gp.kernel_params.lengthscale, gp.kernel_params.vertical_lengthscale,
                    gp.kernel_params.noise_variance]
    return out


@pytest.fixture(scope="module")
def golden(golden_dir):
    with open(os.path.join(golden_dir, "output.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("key", ["choleksy", "sum", "full", "pred", "losses"])
def test_gpu_matches_golden(results, golden, key):
    got, exp = results[key], golden[key]
    print(f"{key}: max rel err", max_rel(np.concatenate([np.ravel(a) for a in got]) if key in ("choleksy", "sum", "full") else got,
                                         np.concatenate([np.ravel(a) for a in exp]) if key in ("choleksy", "sum", "full") else exp))
    if key == "choleksy":
        assert [len(t) for t in got] == [len(t) for t in exp]
        for a, b in zip(got, exp):
            assert np.all(within_rel(a, b, TOL)), max_rel(a, b)
    elif key in ("sum", "full"):
        for a, b in zip(got, exp):
            assert np.all(within_rel(a, b, TOL)), max_rel(a, b)
    else:
        assert np.all(within_rel(got, exp, TOL)), max_rel(got, exp)
