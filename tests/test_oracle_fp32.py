"""The oracle's FP32 comparator of the mixed-precision path (a restatement of
core/src/cpu/adapter_cblas_fp32.cpp:14-100 composed with the FP64 path):
level 1 computes the Cholesky's SYRK/GEMM trailing products through the FP32
adapter, level 2 every factor op; alpha is refined in FP64.

Pins: with one tile there is no trailing product, so level 1 reduces to the
FP64 oracle exactly (variance and loss bit for bit; the mean within 1e-13,
the refinement's FP64 correction); the FP32 error of levels 1 and 2 is of
single-precision size (1e-10 .. 1e-5 normwise) and the variance error shrinks when the
covariance is better conditioned (noise 1.0 against 0.1); the refined mean
stays within 1e-12 of FP64 at every level."""
import numpy as np
import pytest

from conftest import norm_rel


def _run(oracle, T, b, r, noise, m=256):
    from gprat_amd import datagen
    x, y, z = datagen.gp_problem(T * b, m, r)
    g = oracle.GP(list(x), list(y), T, b, r, [1.0, 1.0, noise], [True, True, True])
    mt, mb = oracle.compute_test_tiles(m, T, b)
    om, ov = g.predict_with_uncertainty(list(z), mt, mb)
    ol = g.compute_loss()
    out = {}
    for level in (0, 1, 2):
        mm, vv, ll, steps = g.predict_with_uncertainty_reduced(list(z), mt, mb, level)
        out[level] = (norm_rel(mm, om), norm_rel(vv, ov), abs(ll - ol), steps, vv, ov, ll, ol)
    return out


def test_one_tile_reduces_to_fp64(oracle):
    out = _run(oracle, 1, 512, 8, 0.1)
    em, ev, el, steps, vv, ov, ll, ol = out[1]
    assert np.array_equal(np.asarray(vv), np.asarray(ov)) and ll == ol
    assert em < 1e-13 and 1 <= steps <= 30
    # level 0 (FP64 factor) likewise, at any tile count
    em0, ev0, el0, *_ = _run(oracle, 4, 128, 8, 0.1)[0]
    assert ev0 == 0.0 and el0 < 1e-15 and em0 < 1e-13


@pytest.mark.parametrize("T,b", [(4, 128), (8, 128)])
def test_fp32_error_size_and_conditioning(oracle, T, b):
    ill = _run(oracle, T, b, 8, 0.1)
    well = _run(oracle, T, b, 8, 1.0)
    for level in (1, 2):
        em, ev, el, steps, *_ = ill[level]
        print(f"T={T} b={b} level {level}: mean {em:.1e} var {ev:.1e} loss {el:.1e} ({steps} refinement steps); "
              f"noise 1.0: var {well[level][1]:.1e} loss {well[level][2]:.1e}")
        assert 1e-10 < ev < 1e-5 and el < 1e-6
        assert em < 1e-12
        assert well[level][1] < ev
