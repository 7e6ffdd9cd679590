"""CPU checks of the C ABI and the drop-in facade (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gprat_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gprat_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from gprat_amd import capi
    lib = ctypes.CDLL(capi.LIBRARY_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes binding covers the whole header
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_status_codes_and_errors_without_device():
    from gprat_amd import capi
    lib = capi.lib()
    assert lib.gprat_version().decode().startswith("gprat_amd")
    n = ctypes.c_int(-1)
    assert lib.gprat_device_count(ctypes.byref(n)) == capi.OK
    if n.value == 0:
        h = ctypes.c_void_p()
        assert lib.gprat_context_create(0, 1, ctypes.byref(h)) == capi.ERR_NO_DEVICE
        assert "device" in lib.gprat_last_error().decode().lower()
    # null context is an argument error on every entry point that takes one
    x = np.zeros(16)
    t = np.array([1.0, 1.0, 0.1])
    out = np.zeros(64)
    st = lib.gprat_cholesky(None, x.ctypes.data_as(capi._D), x.size, t.ctypes.data_as(capi._D), 2, 4, 2,
                            out.ctypes.data_as(capi._D))
    assert st == capi.ERR_ARGUMENT
    assert "context" in lib.gprat_last_error().decode()
    loss = ctypes.c_double()
    st = lib.gprat_compute_loss(None, x.ctypes.data_as(capi._D), x.size, x.ctypes.data_as(capi._D), x.size,
                                t.ctypes.data_as(capi._D), 2, 4, 2, ctypes.byref(loss))
    assert st == capi.ERR_ARGUMENT
    assert lib.gprat_model_release(None) == capi.OK


def test_server_arguments_without_device():
    from gprat_amd import capi
    lib = capi.lib()
    h = ctypes.c_void_p()
    bad = capi.ServerConfig(0, 0, 0, capi.PRECISION_MIXED, 2048, 256)  # zero slots
    assert lib.gprat_server_create(ctypes.byref(bad), ctypes.byref(h)) == capi.ERR_ARGUMENT
    assert "slots" in lib.gprat_last_error().decode()
    n = ctypes.c_int(-1)
    assert lib.gprat_device_count(ctypes.byref(n)) == capi.OK
    if n.value == 0:
        cfg = capi.ServerConfig(0, 4, 0, capi.PRECISION_MIXED, 2048, 256)
        assert lib.gprat_server_create(ctypes.byref(cfg), ctypes.byref(h)) == capi.ERR_NO_DEVICE
    r = capi.RequestResult()
    assert lib.gprat_server_wait(None, 1, 0.0, ctypes.byref(r)) == capi.ERR_ARGUMENT
    assert lib.gprat_server_destroy(None) == capi.ERR_ARGUMENT
    t = ctypes.c_longlong()
    assert lib.gprat_server_submit_fit(None, None, ctypes.byref(t)) == capi.ERR_ARGUMENT


def test_facade_api_surface(oracle):
    import gprat_amd as gprat
    for name in ["GP", "GP_data", "AdamParams", "SEKParams", "compute_train_tiles", "compute_train_tile_size",
                 "compute_test_tiles", "print_vector", "start_hpx", "resume_hpx", "suspend_hpx", "stop_hpx",
                 "compiled_with_cuda", "print_available_gpus", "gpu_count"]:
        assert hasattr(gprat, name), name
    for args in [(1024, 4, 256), (1000, 4, 256), (3, 4, 256), (4096, 32, 512), (1, 1, 1)]:
        assert gprat.compute_test_tiles(*args) == oracle.compute_test_tiles(*args)
    assert gprat.compute_train_tiles(1024, 256) == 4
    assert gprat.compute_train_tile_size(1000, 3) == 333
    with pytest.raises(RuntimeError):
        gprat.compute_train_tiles(10, 0)
    a = gprat.AdamParams(0.1, 0.9, 0.999, 1e-8, 3)
    assert a.opt_iter == 3 and "learning_rate=0.1" in repr(a)
    gprat.start_hpx([], 4)
    gprat.stop_hpx()


def test_facade_gp_object(golden_dir):
    import gprat_amd as gprat
    d = gprat.GP_data(os.path.join(golden_dir, "training_input.txt"), 128, 8)
    assert len(d.data) == 135 and d.data[:7] == [0.0] * 7
    with pytest.raises(RuntimeError):
        gprat.GP_data(os.path.join(golden_dir, "missing.txt"), 8, 8)
    gp = gprat.GP(d.data, d.data, 4, 32, 8, [1.0, 1.0, 0.1], [True, True, True], 0, 1)
    assert "lengthscale=1.000000000000" in repr(gp) and "GPU (HIP)" in repr(gp)
    gp.kernel_params.lengthscale = 2.0
    assert gp.kernel_params.lengthscale == 2.0
    assert gp.get_input_data() == d.data
    with pytest.raises(ValueError):
        gprat.GP(d.data, d.data, 4, 32, 8, [1.0, 1.0], [True, True, True])


def test_cpu_only_host_fails_loudly():
    import gprat_amd as gprat
    if gprat.gpu_count() > 0:
        pytest.skip("a GPU is present")
    gp = gprat.GP([0.0] * 20, [0.0] * 20, 2, 8, 4, [1.0, 1.0, 0.1], [True, True, True])
    with pytest.raises(RuntimeError, match="device"):
        gp.compute_loss()
