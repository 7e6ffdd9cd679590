"""The host runtime's life cycle (core/src/utils_c.cpp:123-133 semantics):
start records --hpx:threads, suspend blocks the GP API until resume, stop
ends it, out-of-order calls are refused, and start works again after stop
(the example harness starts and stops once per measurement)."""
import pytest


def test_runtime_life_cycle():
    import gprat_amd as gprat
    from gprat_amd import datagen
    x, y, z = datagen.gp_problem(64, 4, 4)
    gp = gprat.GP(list(x), list(y), 1, 64, 4, [1.0, 1.0, 0.1], [True, True, True])  # host target: no device needed
    assert gprat.runtime_state() in ("not started", "stopped")
    gprat.start_hpx([], 3)
    assert gprat.runtime_state() == "running" and gprat.host_threads() == 3
    with pytest.raises(RuntimeError):
        gprat.start_hpx([], 2)  # already running
    gprat.suspend_hpx()
    assert gprat.runtime_state() == "suspended"
    with pytest.raises(RuntimeError, match="suspended"):
        gp.predict(list(z), 1, 4)
    with pytest.raises(RuntimeError):
        gprat.suspend_hpx()
    gprat.resume_hpx()
    with pytest.raises(RuntimeError, match="GPU"):
        gp.predict(list(z), 1, 4)  # runnable again: this GP has no GPU target
    gprat.stop_hpx()
    assert gprat.runtime_state() == "stopped"
    with pytest.raises(RuntimeError):
        gprat.stop_hpx()
    gprat.start_hpx([], 2)
    assert gprat.host_threads() == 2
    gprat.stop_hpx()
