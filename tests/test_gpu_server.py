"""Request loop with continuous batching: a request's result is the same
alone and in a concurrent mix (bit-for-bit, every kernel reduces in a fixed
order), small fits share batched launches, eviction under a tight HBM
budget rebuilds factors with identical results, and failures are reported
per request."""
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
    # the four small fits went through batched launches (how many shared one
    # depends on arrival timing)
    assert stats["batched_fits"] == 4 and 1 <= stats["batch_launches"] <= 4
    worst = 0.0
    for key, r in ref.items():
        g = got[key]
        assert r["status"] == 0 and g["status"] == 0, (key, r["error"], g["error"])
        if key[0] == "fit":
            assert g["loss"] == r["loss"], key
        else:
            worst = max(worst, norm_rel(g["mean"], r["mean"]), norm_rel(g["var"], r["var"]))
            assert np.array_equal(g["mean"], r["mean"]), key
            assert np.array_equal(g["var"], r["var"]), key
    print(f"precision {precision}: max difference alone vs mix {worst:.1e}; {stats}")


def test_server_matches_resident_models():
    """FP64 server results equal the plain resident-model path (single-GP
    fits) and the batched API (small fits) to 1e-10."""
    from gprat_amd import capi
    ctx = capi.Context(0, 1)
    try:
        with capi.Server(0, slots=2, precision=capi.PRECISION_FP64) as srv:
            out = _run_mix(srv, range(len(MIX)), waves=True)
        for i, (n, r, theta, m) in enumerate(MIX):
            x, y, z = _data(i, n, r, m)
            if n <= 2048:
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
