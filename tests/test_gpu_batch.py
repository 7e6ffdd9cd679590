"""Batched many-GP fit against single-GP fits and the CPU oracle: every GP's
loss, mean and variance agree to 1e-10. The two execution paths (a
persistent kernel running whole GPs, and level-synchronous launches over
the 64x64 blocks of all large GPs, replayed from a captured HIP graph when
a batch composition repeats) give bit-identical results."""
import os

import numpy as np
import pytest

from conftest import norm_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gprat_amd import capi
    c = capi.Context(0, 1)
    yield c
    c.close()


def _problems(count, seed, n_lo=64, n_hi=700):
    from gprat_amd import datagen
    rng = np.random.default_rng(seed)
    probs = []
    for g in range(count):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        r = int(rng.choice([4, 8, 16]))
        theta = (float(np.exp(rng.uniform(np.log(0.3), np.log(3.0)))),
                 float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))),
                 float(np.exp(rng.uniform(np.log(0.01), np.log(0.3)))))
        x, y, z = datagen.gp_problem(n, 50, r, seed_train=1000 + 2 * g, seed_test=1001 + 2 * g)
        probs.append((x, y, n, r, theta, z))
    return probs


def test_batch_matches_single_fits(ctx):
    from gprat_amd import capi
    probs = _problems(48, seed=5)
    batch = capi.Batch(ctx, [(x, y, n, r, t) for x, y, n, r, t, _ in probs])
    try:
        losses = batch.losses()
        worst = 0.0
        for g, (x, y, n, r, t, z) in enumerate(probs):
            single = capi.Model(ctx, x, y, n, r, t)
            sm, sv = single.predict(z, 50)
            bm, bv = batch.predict(g, z, 50)
            sl = single.loss()
            single.release()
            worst = max(worst, norm_rel(bm, sm), norm_rel(bv, sv), abs(losses[g] - sl) / abs(sl))
            assert norm_rel(bm, sm) < 1e-10, g
            assert norm_rel(bv, sv) < 1e-10, g
            assert abs(losses[g] - sl) <= 1e-10 * abs(sl), g
        print(f"48 GPs, worst relative difference to single fits {worst:.3e}")
    finally:
        batch.release()


def test_largest_batched_gps_match_single_fits(ctx):
    """The top of the batched workload's size range (n up to 2048: the
    64-wide level path with its carried diagonal sums, up to 32 block
    columns) against single-GP fits: mean, variance and loss to 1e-10."""
    from gprat_amd import capi, datagen
    sizes = [(1537, 4, (0.4, 1.7, 0.02)), (1800, 16, (2.5, 0.6, 0.25)), (1999, 8, (1.0, 1.0, 0.1)), (2048, 8, (3.0, 2.0, 0.01))]
    probs = []
    for g, (n, r, t) in enumerate(sizes):
        x, y, z = datagen.gp_problem(n, 64, r, seed_train=500 + 2 * g, seed_test=501 + 2 * g)
        probs.append((x, y, n, r, t, z))
    batch = capi.Batch(ctx, [(x, y, n, r, t) for x, y, n, r, t, _ in probs])
    try:
        losses = batch.losses()
        for g, (x, y, n, r, t, z) in enumerate(probs):
            single = capi.Model(ctx, x, y, n, r, t)
            sm, sv = single.predict(z, 64)
            sl = single.loss()
            single.release()
            bm, bv = batch.predict(g, z, 64)
            print(f"n={n}: mean {norm_rel(bm, sm):.2e} var {norm_rel(bv, sv):.2e} loss {abs(losses[g] - sl) / abs(sl):.2e}")
            assert norm_rel(bm, sm) < 1e-10, g
            assert norm_rel(bv, sv) < 1e-10, g
            assert abs(losses[g] - sl) <= 1e-10 * abs(sl), g
    finally:
        batch.release()


def test_batch_matches_oracle(ctx, oracle):
    from gprat_amd import capi
    probs = _problems(6, seed=9, n_lo=100, n_hi=400)
    batch = capi.Batch(ctx, [(x, y, n, r, t) for x, y, n, r, t, _ in probs])
    try:
        losses = batch.losses()
        for g, (x, y, n, r, t, z) in enumerate(probs):
            gp = oracle.GP(list(x), list(y), 1, n, r, list(t), [True] * 3)
            om, ov = gp.predict_with_uncertainty(list(z), 1, 50)
            ol = gp.compute_loss()
            bm, bv = batch.predict(g, z, 50)
            assert norm_rel(bm, om) < 1e-10 and norm_rel(bv, ov) < 1e-10, g
            assert abs(losses[g] - ol) <= 1e-10 * abs(ol), g
    finally:
        batch.release()


def test_batch_reports_not_spd(ctx):
    from gprat_amd import capi
    probs = _problems(3, seed=2)
    bad = [(x, y, n, r, t) for x, y, n, r, t, _ in probs]
    x, y, n, r, t = bad[1]
    bad[1] = (x, y, n, r, (t[0], t[1], -2.0 * t[1]))
    with pytest.raises(capi.GpratError) as e:
        capi.Batch(ctx, bad)
    assert e.value.status == capi.ERR_NOT_SPD and "GP 1" in str(e.value)


@pytest.fixture
def force(monkeypatch):
    """Set the batched fit's routing overrides for one test."""
    def set_(path=None, mixed=None):
        for k, v in (("GPRAT_BATCH_PATH", path), ("GPRAT_BATCH_MIXED", mixed)):
            if v is None:
                monkeypatch.delenv(k, raising=False)
            else:
                monkeypatch.setenv(k, v)
    return set_


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_paths_and_graph_replays_are_bit_identical(ctx, force, mixed):
    """The same batch through the persistent kernel and the level path
    (first direct launches, then the capture on the composition's second
    fit, then replays), with other device work and synchronisations between
    the fits: every replay equals the eager result bit for bit. Mixed
    precision: the level path spreads alpha's refinement over launches (the
    residual per 64-row block), the persistent kernel runs it in one
    workgroup per GP; the refinement steps taken agree too."""
    from gprat_amd import capi
    probs = _problems(24, seed=21, n_lo=64, n_hi=1500)
    fits = [(x, y, n, r, t) for x, y, n, r, t, _ in probs]
    picks = (0, 7, 23)
    precision = capi.PRECISION_MIXED if mixed else capi.PRECISION_FP64

    def run():
        b = capi.Batch(ctx, fits, precision=precision)
        try:
            return b.losses().copy(), b.refinement().copy(), [b.predict(g, probs[g][5], 50) for g in picks]
        finally:
            b.release()

    force(path="persistent", mixed="requested" if mixed else None)
    ref_loss, ref_steps, ref_pred = run()
    if mixed:
        assert np.all(ref_steps >= 1)
    force(path="level", mixed="requested" if mixed else None)
    x, y, n, r, t, z = probs[3]
    for rep in range(4):
        loss, steps, pred = run()
        assert np.array_equal(loss, ref_loss), rep
        assert np.array_equal(steps, ref_steps), rep
        for (m1, v1), (m2, v2) in zip(pred, ref_pred):
            assert np.array_equal(m1, m2) and np.array_equal(v1, v2), rep
        # eager work of the program between the fits
        mdl = capi.Model(ctx, x, y, n, r, t)
        mdl.predict(z, 50)
        mdl.release()
        ctx.synchronize()


def test_tuned_mixed_batch_runs_fp64_where_mixed_is_slower(ctx, force):
    """With the committed tuning table (mixed precision measured slower than
    FP64 for every size class of the batched fit), a batch that asks for
    mixed precision runs every GP in FP64: no refinement steps, results equal
    to the FP64 batch."""
    from gprat_amd import capi
    force()
    probs = _problems(12, seed=31, n_lo=64, n_hi=2048)
    fits = [(x, y, n, r, t) for x, y, n, r, t, _ in probs]
    b64 = capi.Batch(ctx, fits)
    bmx = capi.Batch(ctx, fits, precision=capi.PRECISION_MIXED)
    try:
        assert np.all(bmx.refinement() == 0)
        assert np.array_equal(bmx.losses(), b64.losses())
        m1, v1 = bmx.predict(5, probs[5][5], 50)
        m2, v2 = b64.predict(5, probs[5][5], 50)
        assert np.array_equal(m1, m2) and np.array_equal(v1, v2)
    finally:
        b64.release()
        bmx.release()


def test_mixed_batch_vs_fp64_batch(ctx, force):
    """Mixed precision batch (FP32 block products, FP64 refinement of alpha
    and of the variance) against the FP64 batch on the same 48 GPs: mean
    1e-8 normwise (the refined alpha; the bar of the mixed-precision
    configuration). The variance and the loss (its
    log-determinant comes from the mixed factor) carry the factor's own
    accuracy: every off-diagonal block product rounds its operands to FP32,
    and these GPs reach condition numbers of ~4e5 (noise down to 0.01,
    lengthscale up to 3), so their bars are 1e-3 and 1e-4 (measured
    1.4e-4 and 2.0e-5). Every GP took at least one refinement step."""
    from gprat_amd import capi
    force(mixed="requested")  # mixed precision for every GP, whatever the tuning table says
    probs = _problems(48, seed=5, n_lo=64, n_hi=2048)
    fits = [(x, y, n, r, t) for x, y, n, r, t, _ in probs]
    b64 = capi.Batch(ctx, fits)
    b32 = capi.Batch(ctx, fits, precision=capi.PRECISION_MIXED)
    try:
        l64, l32 = b64.losses(), b32.losses()
        steps = b32.refinement()
        assert np.all(steps >= 1) and np.all(steps <= 10)
        assert np.all(b64.refinement() == 0)
        worst_mean = worst_var = worst_loss = 0.0
        for g, (x, y, n, r, t, z) in enumerate(probs):
            m64, v64 = b64.predict(g, z, 50)
            m32, v32 = b32.predict(g, z, 50)
            worst_mean = max(worst_mean, norm_rel(m32, m64))
            worst_var = max(worst_var, norm_rel(v32, v64))
            worst_loss = max(worst_loss, abs(l32[g] - l64[g]) / abs(l64[g]))
        print(f"mixed batch vs FP64 batch: mean {worst_mean:.2e} var {worst_var:.2e} loss {worst_loss:.2e}, "
              f"refinement steps {steps.min()}..{steps.max()}")
        assert worst_mean < 1e-8
        assert worst_var < 1e-3
        assert worst_loss < 1e-4
    finally:
        b64.release()
        b32.release()
