"""The entry point that allocates nothing (gprat_model_predict_mean_noalloc)
against gprat_model_predict's mean: bit-for-bit equal (same kernels, same
reduction splits), on two calls in a row with the second smaller, for an
FP64 and a mixed-precision model; a request beyond the fixed scratch fails
with an argument error before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_mean_matches_predict():
    from gprat_amd import capi, datagen
    ctx = capi.Context(0, 1)
    try:
        for n, r, prec in ((4096, 8, capi.PRECISION_FP64), (5000, 16, capi.PRECISION_MIXED)):
            x, y, z = datagen.gp_problem(n, 70000, r)
            mdl = capi.Model(ctx, x, y, n, r, (1.0, 1.0, 0.1), prec)
            try:
                for m in (70000, 37):
                    zz = z[:m + r - 1]
                    want = mdl.predict(zz, m, variance=False)
                    got = mdl.predict_mean_noalloc(zz, m)
                    print(f"n={n} precision={prec} m={m}: max |diff| {np.max(np.abs(got - want)):.1e}")
                    assert np.array_equal(got, want), (n, m)
                big = 400000
                zb = np.zeros(big + r - 1)
                with pytest.raises(capi.GpratError) as e:
                    mdl.predict_mean_noalloc(zb, big)
                assert e.value.status == capi.ERR_ARGUMENT and "scratch" in str(e.value)
            finally:
                mdl.release()
    finally:
        ctx.close()
