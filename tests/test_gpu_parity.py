"""GPU path (through the C ABI) against the CPU oracle on seeded inputs.

Bar: 1e-10 relative for mean, variance and loss (normwise for vectors,
||a-b||_inf / ||b||_inf, since mean entries near zero inflate element-wise
error) and for the Cholesky tiles (normwise per tile); the gradient is
compared in relative terms at 1e-8 (it is a difference of two O(n) traces).
"""
import numpy as np
import pytest

from conftest import norm_rel, max_rel

pytestmark = pytest.mark.gpu

TOL = 1e-10

# (n_tiles, n_tile_size, n_test, n_reg, theta)
CASES = [
    (4, 32, 128, 8, (1.0, 1.0, 0.1)),        # the golden-test shape
    (4, 256, 1024, 8, (1.0, 1.0, 0.1)),
    (3, 100, 70, 8, (1.3, 0.7, 0.05)),       # ragged: n=300, n_pad=384, m=69
    (5, 77, 1, 1, (0.6, 2.0, 0.2)),          # n_reg=1, single test point
    (2, 320, 257, 16, (2.0, 1.5, 0.01)),     # n_reg=16, n=640 (partial outer block)
]


@pytest.fixture(scope="module")
def ctx():
    from gprat_amd import capi
    c = capi.Context(0, 1)
    yield c
    c.close()


def _problem(T, b, m_req, r, seed=0):
    from gprat_amd import datagen
    x, y, z = datagen.gp_problem(T * b, max(m_req, 1), r, seed_train=seed, seed_test=seed + 1)
    return x, y, z


def _tiles(oracle, m_req, T, b):
    if m_req == 1:
        return 1, 1
    return oracle.compute_test_tiles(m_req, T, b)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"T{c[0]}b{c[1]}m{c[2]}r{c[3]}")
def test_predict_with_uncertainty(ctx, oracle, case):
    T, b, m_req, r, theta = case
    x, y, z = _problem(T, b, m_req, r)
    mt, mb = _tiles(oracle, m_req, T, b)
    gp = oracle.GP(list(x), list(y), T, b, r, list(theta), [True] * 3)
    om, ov = gp.predict_with_uncertainty(list(z), mt, mb)
    gm, gv = ctx.predict(x, y, z, theta, T, b, mt, mb, r, variance=True)
    print(f"case {case}: mean normrel {norm_rel(gm, om):.3e} var normrel {norm_rel(gv, ov):.3e} "
          f"mean maxrel {max_rel(gm, om):.3e}")
    assert norm_rel(gm, om) < TOL
    assert norm_rel(gv, ov) < TOL
    fm, fv = ctx.predict(x, y, z, theta, T, b, mt, mb, r, full_cov=True)
    ofm, ofv = gp.predict_with_full_cov(list(z), mt, mb)
    assert norm_rel(fm, ofm) < TOL and norm_rel(fv, ofv) < TOL
    pm = ctx.predict(x, y, z, theta, T, b, mt, mb, r)
    # predict factors K alone and forms K* alpha; predict_with_uncertainty
    # factors the test rows and the targets below K (the same kernels; the
    # larger trailing updates may be split over workgroups differently) and
    # forms the mean as (K* L^-T)(L^-1 y), so they agree to rounding
    assert norm_rel(pm, gm) < 1e-11


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"T{c[0]}b{c[1]}m{c[2]}r{c[3]}")
def test_loss_and_cholesky(ctx, oracle, case):
    T, b, m_req, r, theta = case
    x, y, _ = _problem(T, b, m_req, r)
    gp = oracle.GP(list(x), list(y), T, b, r, list(theta), [True] * 3)
    ol = gp.compute_loss()
    gl = ctx.compute_loss(x, y, theta, T, b, r)
    print(f"case {case}: loss {gl!r} oracle {ol!r} rel {abs(gl - ol) / abs(ol):.3e}")
    assert abs(gl - ol) <= TOL * abs(ol)
    tiles = ctx.cholesky(x, theta, T, b, r)
    ot = gp.cholesky()
    k = 0
    worst = 0.0
    for i in range(T):
        for j in range(i + 1):
            ref = np.array(ot[i * T + j]).reshape(b, b)
            worst = max(worst, norm_rel(tiles[k], ref))
            k += 1
    print(f"case {case}: cholesky tiles worst normrel {worst:.3e}")
    assert worst < TOL


@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: f"T{c[0]}b{c[1]}m{c[2]}r{c[3]}")
def test_gradient_and_adam(ctx, oracle, case):
    T, b, m_req, r, theta = case
    x, y, _ = _problem(T, b, m_req, r)
    # oracle gradient recovered from the first Adam moment after one step
    gp = oracle.GP(list(x), list(y), T, b, r, list(theta), [True] * 3)
    ol = gp.optimize_step(oracle.AdamParams(0.1, 0.9, 0.999, 1e-8, 1), 0)
    og = np.array(gp.adam_moments[0]) / (1 - 0.9)
    gl, gg = ctx.loss_gradient(x, y, theta, T, b, r)
    print(f"case {case}: grad {gg} oracle {og} loss rel {abs(gl - ol) / abs(ol):.3e}")
    assert abs(gl - ol) <= TOL * abs(ol)
    assert np.allclose(gg, og, rtol=1e-8, atol=1e-12)


def test_optimize_three_steps_matches_oracle(oracle):
    import gprat_amd as gprat
    T, b, r = 4, 64, 8
    x, y, _ = _problem(T, b, 1, r, seed=7)
    adam_o = oracle.AdamParams(0.1, 0.9, 0.999, 1e-8, 3)
    go = oracle.GP(list(x), list(y), T, b, r, [1.0, 1.0, 0.1], [True, True, True])
    lo = go.optimize(adam_o)
    gg = gprat.GP(list(x), list(y), T, b, r, [1.0, 1.0, 0.1], [True, True, True], 0, 1)
    lg = gg.optimize(gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, 3))
    th = [gg.kernel_params.lengthscale, gg.kernel_params.vertical_lengthscale, gg.kernel_params.noise_variance]
    print("losses", lg, lo, "theta", th, go.hyperparameters)
    assert max_rel(lg, lo) < TOL
    assert max_rel(th, go.hyperparameters) < 1e-9


def test_partial_trainable(oracle):
    import gprat_amd as gprat
    T, b, r = 2, 64, 4
    x, y, _ = _problem(T, b, 1, r, seed=3)
    for trainable in ([True, False, True], [False, True, False], [False, False, False]):
        go = oracle.GP(list(x), list(y), T, b, r, [0.7, 1.2, 0.3], trainable)
        lo = go.optimize(oracle.AdamParams(0.05, 0.9, 0.999, 1e-8, 2))
        gg = gprat.GP(list(x), list(y), T, b, r, [0.7, 1.2, 0.3], trainable, 0, 1)
        lg = gg.optimize(gprat.AdamParams(0.05, 0.9, 0.999, 1e-8, 2))
        th = [gg.kernel_params.lengthscale, gg.kernel_params.vertical_lengthscale, gg.kernel_params.noise_variance]
        assert max_rel(lg, lo) < TOL
        assert max_rel(th, go.hyperparameters) < 1e-9


def test_empty_test_set(ctx):
    # compute_test_tiles(3, 4, 256) = (4, 0): the reference returns empty vectors
    x, y, z = _problem(4, 64, 3, 8)
    m, v = ctx.predict(x, y, z, (1.0, 1.0, 0.1), 4, 64, 4, 0, 8, variance=True)
    assert m.size == 0 and v.size == 0


def test_not_spd_is_reported(ctx):
    from gprat_amd import capi
    x, y, z = _problem(2, 64, 8, 4)
    with pytest.raises(capi.GpratError) as e:
        ctx.compute_loss(x, y, (1.0, 1.0, -5.0), 2, 64, 4)  # negative noise: K not SPD
    assert e.value.status == capi.ERR_NOT_SPD


def test_bad_arguments(ctx):
    from gprat_amd import capi
    x, y, z = _problem(2, 64, 8, 4)
    with pytest.raises(capi.GpratError) as e:
        ctx.compute_loss(x[:100], y, (1.0, 1.0, 0.1), 2, 64, 4)  # input shorter than n + n_reg - 1
    assert e.value.status == capi.ERR_ARGUMENT
