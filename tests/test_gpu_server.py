"""Request loop with continuous batching: a request's result is the same
alone and in a concurrent mix, small fits share batched launches when that
pays, eviction under a tight HBM budget rebuilds factors with identical
results, and failures are reported per request.

"The same" is bit-for-bit when the request ran on the same path alone and in
the mix (every kernel reduces in a fixed order). A small fit (n <= 2048) is
routed by load: alone it takes the single-GP path, in a mix it may share a
batched launch; both paths are FP64, so such a result agrees to 1e-10
(normwise for mean and variance, relative for the loss)."""
import numpy as np
import pytest

from conftest import norm_rel

pytestmark = pytest.mark.gpu

# (n, n_reg, theta, m); n <= 2048 goes through the batched kernel
MIX = [
    (96, 8, (1.0, 1.0, 0.1), 40),
    (300, 4, (0.7, 1.3, 0.05), 500),
    (777, 8, (1.5, 0.8, 0.2), 1),
    (2048, 16, (1.0, 1.0, 0.1), 257),
    (2049, 8, (1.0, 1.0, 0.05), 1000),
    (5000, 8, (1.2, 1.0, 0.1), 333),
    (9000, 8, (1.0, 1.0, 0.05), 2048),
]


def _data(i, n, r, m):
    from gprat_amd import datagen
    return datagen.gp_problem(n, m, r, seed_train=500 + 2 * i, seed_test=501 + 2 * i)


def _run_mix(server, order, waves):
    """Submit fits and predicts for the MIX entries in `order`; with `waves`
    all requests are in flight together, otherwise each waits for the last."""
    fits, preds, out = {}, {}, {}
    for i in order:
        n, r, theta, m = MIX[i]
        x, y, z = _data(i, n, r, m)
        fits[i] = server.submit_fit(x, y, n, r, theta)
        if not waves:
            out[("fit", i)] = server.wait(fits[i])
        preds[i] = server.submit_predict(fits[i], z, m)
        if not waves:
            out[("pred", i)] = server.wait(preds[i])
    if waves:
        for i in order:
            out[("fit", i)] = server.wait(fits[i])
            out[("pred", i)] = server.wait(preds[i])
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_same_alone_and_in_the_mix(precision):
    from gprat_amd import capi
    with capi.Server(0, slots=4, precision=precision) as alone:
        ref = {}
        for i in range(len(MIX)):
            ref.update(_run_mix(alone, [i], waves=False))
    order = [3, 0, 6, 1, 5, 2, 4]
    with capi.Server(0, slots=4, precision=precision) as mixed:
        got = _run_mix(mixed, order, waves=True)
        stats = mixed.stats()
    assert stats["completed"] == 2 * len(MIX) and stats["failed"] == 0
    # alone, every fit ran by itself: a lone small fit takes the single-GP path
    assert all(ref[("fit", i)]["batched"] == 0 for i in range(len(MIX)))
    assert stats["batched_fits"] <= 4
    worst = 0.0
    for key, r in ref.items():
        g = got[key]
        assert r["status"] == 0 and g["status"] == 0, (key, r["error"], g["error"])
        same_path = (ref[("fit", key[1])]["batched"] > 0) == (got[("fit", key[1])]["batched"] > 0)
        if key[0] == "fit":
            if same_path:
                assert g["loss"] == r["loss"], key
            else:
                assert abs(g["loss"] - r["loss"]) <= 1e-10 * abs(r["loss"]), key
        else:
            worst = max(worst, norm_rel(g["mean"], r["mean"]), norm_rel(g["var"], r["var"]))
            if same_path:
                assert np.array_equal(g["mean"], r["mean"]), key
                assert np.array_equal(g["var"], r["var"]), key
            else:
                assert norm_rel(g["mean"], r["mean"]) <= 1e-10 and norm_rel(g["var"], r["var"]) <= 1e-10, key
    print(f"precision {precision}: max difference alone vs mix {worst:.1e}; {stats}")


def test_cached_blocks_move_between_slots():
    """The slots reuse each other's cached blocks (after waiting for the
    block's last use on its former slot): the mix, run again and again in
    other orders on one server with every model released in between, gives
    bit-identical results on the same path and never has to flush the
    caches."""
    from gprat_amd import capi
    orders = [[3, 0, 6, 1, 5, 2, 4], [6, 5, 4, 3, 2, 1, 0], [0, 6, 2, 5, 4, 1, 3], [5, 3, 1, 6, 0, 4, 2]]
    with capi.Server(0, slots=4, precision=capi.PRECISION_MIXED) as srv:
        runs = []
        for order in orders:
            out = _run_mix(srv, order, waves=True)
            for i in order:
                srv.release_model(out[("pred", i)]["model"])
            runs.append(out)
        stats = srv.stats()
    assert stats["failed"] == 0 and stats["cache_flushes"] == 0, stats
    first = runs[0]
    for out in runs[1:]:
        for key, r in first.items():
            g = out[key]
            if (first[("fit", key[1])]["batched"] > 0) != (out[("fit", key[1])]["batched"] > 0):
                continue
            if key[0] == "fit":
                assert g["loss"] == r["loss"], key
            else:
                assert np.array_equal(g["mean"], r["mean"]) and np.array_equal(g["var"], r["var"]), key
    print(f"blocks borrowed between slots: {stats['borrowed_blocks']}; {stats}")


def test_small_fits_waiting_together_share_a_launch():
    """Many small fits queued at once run as batched launches; their
    results agree with the single-GP path to 1e-10."""
    from gprat_amd import capi
    probs = [(_data(200 + k, 300 + 7 * k, 8, 50), 300 + 7 * k) for k in range(48)]
    with capi.Server(0, slots=2, precision=capi.PRECISION_MIXED) as srv:
        ids = [srv.submit_fit(x, y, n, 8, (1.0, 1.0, 0.1)) for (x, y, _), n in probs]
        fits = [srv.wait(t) for t in ids]
        stats = srv.stats()
    assert all(f["status"] == 0 for f in fits)
    assert stats["batched_fits"] >= 24 and stats["batch_launches"] >= 1, stats
    ctx = capi.Context(0, 1)
    try:
        for ((x, y, _), n), f in zip(probs[:6], fits[:6]):
            mdl = capi.Model(ctx, x, y, n, 8, (1.0, 1.0, 0.1), capi.PRECISION_FP64)
            assert abs(f["loss"] - mdl.loss()) <= 1e-10 * abs(mdl.loss())
            mdl.release()
    finally:
        ctx.close()


def test_server_matches_resident_models():
    """FP64 server results equal the plain resident-model path (single-GP
    fits) and the batched API (batched small fits) to 1e-10."""
    from gprat_amd import capi
    ctx = capi.Context(0, 1)
    try:
        with capi.Server(0, slots=2, precision=capi.PRECISION_FP64) as srv:
            out = _run_mix(srv, range(len(MIX)), waves=True)
        for i, (n, r, theta, m) in enumerate(MIX):
            x, y, z = _data(i, n, r, m)
            if out[("fit", i)]["batched"]:
                b = capi.Batch(ctx, [(x, y, n, r, theta)])
                loss = b.losses()[0]
                mean, var = b.predict(0, z, m)
                b.release()
            else:
                mdl = capi.Model(ctx, x, y, n, r, theta, capi.PRECISION_FP64)
                loss = mdl.loss()
                mean, var = mdl.predict(z, m)
                mdl.release()
            assert abs(out[("fit", i)]["loss"] - loss) <= 1e-10 * abs(loss)
            assert norm_rel(out[("pred", i)]["mean"], mean) <= 1e-10
            assert norm_rel(out[("pred", i)]["var"], var) <= 1e-10
    finally:
        ctx.close()


def test_eviction_rebuilds_identical_factors():
    from gprat_amd import capi
    n, r, m = 4096, 8, 700
    theta = (1.0, 1.0, 0.1)
    probs = [_data(40 + k, n, r, m) for k in range(5)]

    def run(budget):
        with capi.Server(0, slots=2, hbm_budget=budget, precision=capi.PRECISION_MIXED) as srv:
            ids = [srv.submit_fit(x, y, n, r, theta) for x, y, _ in probs]
            fits = [srv.wait(t) for t in ids]
            preds = [srv.wait(srv.submit_predict(mid, z, m)) for mid, (_, _, z) in zip(ids, probs)]
            return fits, preds, srv.stats()

    big_fits, big_preds, big_stats = run(0)
    # room for about three resident factors (4096 rows, leading dimension 4160)
    small_fits, small_preds, small_stats = run(int(3.3 * 8 * 4096 * 4160))
    assert big_stats["evictions"] == 0
    assert small_stats["evictions"] >= 2 and small_stats["refits"] >= 2, small_stats
    assert small_stats["peak_bytes"] <= small_stats["budget"]
    assert any(p["refit"] for p in small_preds)
    for a, b in zip(big_fits, small_fits):
        assert a["status"] == 0 and b["status"] == 0 and a["loss"] == b["loss"]
    for a, b in zip(big_preds, small_preds):
        assert a["status"] == 0 and b["status"] == 0, b["error"]
        assert np.array_equal(a["mean"], b["mean"]) and np.array_equal(a["var"], b["var"])


def test_predict_keeps_its_own_least_recently_used_factor():
    """Two resident factors, then a predict on the least recently used one
    under a budget one byte short of what the predict needs beside both:
    admission must evict the other factor, not the one the predict reads
    (which would cost a refit with bytes never reserved)."""
    from gprat_amd import capi
    n, r, m = 4096, 8, 2000
    theta = (1.0, 1.0, 0.1)
    (xa, ya, za), (xb, yb, _) = _data(60, n, r, m), _data(61, n, r, m)

    def run(budget):
        with capi.Server(0, slots=1, hbm_budget=budget, precision=capi.PRECISION_MIXED) as srv:
            a = srv.submit_fit(xa, ya, n, r, theta)
            assert srv.wait(a)["status"] == 0
            b = srv.submit_fit(xb, yb, n, r, theta)
            assert srv.wait(b)["status"] == 0
            p = srv.wait(srv.submit_predict(a, za, m))
            return p, srv.stats()

    p1, s1 = run(0)
    assert s1["evictions"] == 0
    p2, s2 = run(int(s1["peak_bytes"]) - 1)
    print(f"peak with both factors resident {s1['peak_bytes']}; tight budget: {s2}")
    assert p2["status"] == 0, p2["error"]
    assert s2["evictions"] == 1 and s2["refits"] == 0 and not p2["refit"], s2
    assert s2["peak_bytes"] <= s2["budget"]
    assert np.array_equal(p1["mean"], p2["mean"]) and np.array_equal(p1["var"], p2["var"])


def test_failures_are_per_request():
    from gprat_amd import capi
    x, y, z = _data(90, 600, 8, 10)
    x2, y2, z2 = _data(91, 6000, 8, 10)
    with capi.Server(0, slots=2, hbm_budget=256 << 20) as srv:
        with pytest.raises(capi.GpratError) as e:
            srv.submit_predict(12345, z, 10)
        assert e.value.status == capi.ERR_ARGUMENT
        bad = srv.submit_fit(x, y, 600, 8, (1.0, 1.0, -3.0))   # noise < 0: K is not positive definite
        good = srv.submit_fit(x, y, 600, 8, (1.0, 1.0, 0.1))
        bad_pred = srv.submit_predict(bad, z, 10)
        huge = srv.submit_fit(x2, y2, 6000, 8, (1.0, 1.0, 0.1))  # its factor alone exceeds the 256 MiB budget
        assert srv.wait(bad)["status"] == capi.ERR_NOT_SPD
        r = srv.wait(bad_pred)
        assert r["status"] == capi.ERR_NOT_SPD and "failed to fit" in r["error"]
        assert srv.wait(good)["status"] == 0
        assert srv.wait(huge)["status"] == capi.ERR_OUT_OF_MEMORY
        p = srv.wait(srv.submit_predict(good, z, 10))
        assert p["status"] == 0 and np.all(np.isfinite(p["mean"]))
        srv.release_model(good)
        with pytest.raises(capi.GpratError):
            srv.submit_predict(good, z, 10)
        with pytest.raises(capi.GpratError):
            srv.wait(999999)
        with pytest.raises(capi.GpratError):
            srv.wait(bad)  # already retired


def test_serving_slice_at_full_size(oracle):
    """A seeded slice of the serving workload at its full sizes: fits from
    n = 300 to 32768 (n_reg 8, mixed precision with FP64 refinement of alpha
    and of the variance for n > 2048, FP64 below) and predicts with m up to
    4096 against the fitted models, under the 200 GB budget. Each result is
    the same alone and in the concurrent mix (same paths: bit-for-bit), and
    the largest and the smallest requests match the oracle. The largest
    (mixed precision): mean 1e-8 (normwise), variance 1e-6 (normwise; the
    FP64 refinement leaves an error second order in the FP32 solves', 1.3e-7
    measured here: 8-long windows make this covariance worse conditioned than
    the 16-long ones of the mixed-precision configuration), loss 1e-6
    absolute (the mixed-precision factor's log-determinant, unrefined). The
    smallest (FP64): 1e-10."""
    from gprat_amd import capi
    # (n, m, tiles for the oracle)
    slice_ = [(32768, 4096, 32), (300, 300, 3), (12000, 1, 12), (3000, 2500, 6), (1500, 37, 3), (700, 4096, 7)]
    theta = (1.0, 1.0, 0.1)
    data = [_data(300 + i, n, 8, m) for i, (n, m, _) in enumerate(slice_)]

    def run(order, together):
        out = {}
        with capi.Server(0, slots=4, hbm_budget=200 << 30, precision=capi.PRECISION_MIXED) as srv:
            tickets = {}
            for i in order:
                (x, y, z), (n, m, _) = data[i], slice_[i]
                f = srv.submit_fit(x, y, n, 8, theta)
                p = srv.submit_predict(f, z, m)
                tickets[i] = (f, p)
                if not together:
                    out[i] = (srv.wait(f), srv.wait(p))
            for i in order:
                if together:
                    out[i] = (srv.wait(tickets[i][0]), srv.wait(tickets[i][1]))
        return out

    alone = {}
    for i in range(len(slice_)):
        alone.update(run([i], together=False))
    mix = run([2, 0, 5, 1, 4, 3], together=True)
    for i in range(len(slice_)):
        (fa, pa), (fm, pm) = alone[i], mix[i]
        assert fa["status"] == 0 and pa["status"] == 0 and fm["status"] == 0 and pm["status"] == 0, (i, pm["error"])
        assert (fa["batched"] > 0) == (fm["batched"] > 0) or slice_[i][0] <= 2048
        if (fa["batched"] > 0) == (fm["batched"] > 0):
            assert fa["loss"] == fm["loss"], i
            assert np.array_equal(pa["mean"], pm["mean"]) and np.array_equal(pa["var"], pm["var"]), i
        else:
            assert norm_rel(pa["mean"], pm["mean"]) <= 1e-10 and norm_rel(pa["var"], pm["var"]) <= 1e-10, i
    for i in (0, 1):
        (x, y, z), (n, m, T) = data[i], slice_[i]
        b = n // T
        go = oracle.GP(list(x), list(y), T, b, 8, list(theta), [True, True, True])
        mt, mb = oracle.compute_test_tiles(m, T, b)
        assert mt * mb == m
        om, ov = go.predict_with_uncertainty(list(z), mt, mb)
        ol = go.compute_loss()
        f, p = alone[i]
        em, ev, el = norm_rel(p["mean"], om), norm_rel(p["var"], ov), abs(f["loss"] - ol)
        print(f"serving n={n} m={m} vs oracle: mean {em:.2e} var {ev:.2e} loss abs {el:.2e}")
        if n > 2048:
            assert em < 1e-8 and ev < 1e-6 and el < 1e-6
        else:
            assert em < 1e-10 and ev < 1e-10 and el < 1e-10 * abs(ol)
