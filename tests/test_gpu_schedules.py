"""The Cholesky's schedule knobs change only how the work is grouped, never
the result: parity with the oracle (reference cpu:: path,
core/src/cpu/gp_functions.cpp predict_with_uncertainty) under panel widths
that leave a narrower last panel (diagonal-block task schedules of 6 and 4
blocks), one workgroup per panel-stream launch (every task of a launch run
in order by one workgroup), few head workgroups, and the single-stream path
without look-ahead."""
import os

import pytest

from conftest import norm_rel

pytestmark = pytest.mark.gpu

TOL = 1e-10

VARIANTS = [
    {"GPRAT_CHOLESKY_BLOCK": "384"},
    {"GPRAT_CHOLESKY_BLOCK": "640", "GPRAT_TRAILING_RESERVE": "1", "GPRAT_HEAD_WORKGROUPS": "1"},
    {"GPRAT_LOOKAHEAD": "0", "GPRAT_CHOLESKY_BLOCK": "384"},
]


@pytest.fixture(scope="module")
def ctx():
    from gprat_amd import capi
    c = capi.Context(0, 1)
    yield c
    c.close()


@pytest.fixture(scope="module")
def problem(oracle):
    from gprat_amd import datagen
    T, b, m, r, theta = 10, 400, 300, 8, (1.0, 1.0, 0.1)  # n = 4000: n_pad 4096
    x, y, z = datagen.gp_problem(T * b, m, r)
    mt, mb = oracle.compute_test_tiles(m, T, b)
    gp = oracle.GP(list(x), list(y), T, b, r, list(theta), [True, True, True])
    om, ov = gp.predict_with_uncertainty(list(z), mt, mb)
    return (x, y, z, theta, T, b, mt, mb, r), om, ov


@pytest.mark.parametrize("env", VARIANTS, ids=lambda e: ",".join(f"{k[6:]}={v}" for k, v in e.items()))
def test_schedule_variants_match_oracle(ctx, problem, env):
    args, om, ov = problem
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        gm, gv = ctx.predict(*args, variance=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    print(f"{env}: mean {norm_rel(gm, om):.3e} var {norm_rel(gv, ov):.3e}")
    assert norm_rel(gm, om) < TOL
    assert norm_rel(gv, ov) < TOL
