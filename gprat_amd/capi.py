"""ctypes binding of the C ABI in include/gprat_amd.h.

This is how a non-C++ host would bind the backend (the reference's pybind
module is the other route). Every call raises GpratError on a non-zero
status, carrying the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIBRARY_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgprat_amd.so")

OK, ERR_ARGUMENT, ERR_HIP, ERR_OUT_OF_MEMORY, ERR_NOT_SPD, ERR_NO_DEVICE, ERR_INTERNAL, ERR_PENDING = range(8)
PRECISION_FP64, PRECISION_MIXED = 0, 1

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_VP = C.c_void_p
_SZ = C.c_size_t
_LL = C.c_longlong


class ServerConfig(C.Structure):
    """gprat_server_config."""
    _fields_ = [("device", C.c_int), ("slots", C.c_int), ("hbm_budget", _SZ), ("precision", C.c_int),
                ("batch_max_n", C.c_int), ("batch_max", C.c_int)]


class RequestResult(C.Structure):
    """gprat_request_result."""
    _fields_ = [("status", C.c_int), ("kind", C.c_int), ("model", _LL), ("loss", C.c_double),
                ("queue_ms", C.c_double), ("service_ms", C.c_double), ("refit", C.c_int), ("batched", C.c_int),
                ("error", C.c_char * 256)]


class ServerStats(C.Structure):
    """gprat_server_stats."""
    _fields_ = [(k, _LL) for k in ("submitted", "completed", "failed", "fits", "predicts", "batched_fits",
                                   "batch_launches", "evictions", "refits", "resident_models")] + \
               [(k, _SZ) for k in ("resident_bytes", "reserved_bytes", "peak_bytes", "budget")] + \
               [(k, _LL) for k in ("cache_flushes", "borrowed_blocks")]


# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "gprat_last_error": [],
    "gprat_version": [],
    "gprat_device_count": [_I],
    "gprat_device_describe": [C.c_int, C.c_char_p, _SZ],
    "gprat_context_create": [C.c_int, C.c_int, C.POINTER(_VP)],
    "gprat_context_destroy": [_VP],
    "gprat_context_synchronize": [_VP],
    "gprat_device_alloc": [_VP, _SZ, C.POINTER(_VP)],
    "gprat_device_free": [_VP, _VP],
    "gprat_memcpy": [_VP, _VP, _VP, _SZ],
    "gprat_cholesky": [_VP, _D, _SZ, _D, C.c_int, C.c_int, C.c_int, _D],
    "gprat_predict": [_VP, _D, _SZ, _D, _SZ, _D, _SZ, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _D],
    "gprat_predict_with_uncertainty": [_VP, _D, _SZ, _D, _SZ, _D, _SZ, _D, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, _D, _D],
    "gprat_predict_with_full_cov": [_VP, _D, _SZ, _D, _SZ, _D, _SZ, _D, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, _D, _D],
    "gprat_predict_full_posterior": [_VP, _D, _SZ, _D, _SZ, _D, _SZ, _D, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _D, _D],
    "gprat_compute_loss": [_VP, _D, _SZ, _D, _SZ, _D, C.c_int, C.c_int, C.c_int, _D],
    "gprat_loss_gradient": [_VP, _D, _SZ, _D, _SZ, _D, C.c_int, C.c_int, C.c_int, _I, _D, _D],
    "gprat_model_fit": [_VP, _D, _SZ, _D, _SZ, C.c_int, C.c_int, _D, C.c_int, C.POINTER(_VP)],
    "gprat_model_predict": [_VP, _D, _SZ, C.c_int, _D, _D],
    "gprat_model_predict_full_cov": [_VP, _D, _SZ, C.c_int, _D, _D],
    "gprat_model_predict_mean_noalloc": [_VP, _D, _SZ, C.c_int, _D],
    "gprat_model_loss": [_VP, _D],
    "gprat_model_gradient": [_VP, _I, _D, _D],
    "gprat_model_bytes": [_VP, C.POINTER(_SZ)],
    "gprat_model_refinement": [_VP, _I, _D],
    "gprat_model_set_variance_refinement": [_VP, C.c_int],
    "gprat_model_variance_correction": [_VP, _D],
    "gprat_model_release": [_VP],
    "gprat_batch_fit": [_VP, C.c_int, C.c_void_p, C.c_int, C.POINTER(_VP)],
    "gprat_batch_losses": [_VP, _D],
    "gprat_batch_refinement": [_VP, _I],
    "gprat_batch_predict": [_VP, C.c_int, _D, _SZ, C.c_int, _D, _D],
    "gprat_batch_release": [_VP],
    "gprat_server_create": [C.POINTER(ServerConfig), C.POINTER(_VP)],
    "gprat_server_destroy": [_VP],
    "gprat_server_submit_fit": [_VP, C.c_void_p, C.POINTER(_LL)],
    "gprat_server_submit_predict": [_VP, _LL, _D, _SZ, C.c_int, _D, _D, C.POINTER(_LL)],
    "gprat_server_wait": [_VP, _LL, C.c_double, C.POINTER(RequestResult)],
    "gprat_server_release_model": [_VP, _LL],
    "gprat_server_get_stats": [_VP, C.POINTER(ServerStats)],
    "gprat_context_set_profiling": [_VP, C.c_int],
    "gprat_device_calibrate": [_VP, _SZ, _D, _D, _D],
    "gprat_context_kernel_stats": [_VP, C.c_char_p, C.POINTER(C.c_long), _D],
}
_RESTYPE = {"gprat_last_error": C.c_char_p, "gprat_version": C.c_char_p}


class GpDesc(C.Structure):
    """gprat_gp_desc: one GP of a batch (host pointers)."""
    _fields_ = [("train_in", _D), ("train_in_len", _SZ), ("train_out", _D), ("train_out_len", _SZ),
                ("n", C.c_int), ("n_reg", C.c_int), ("theta", C.c_double * 3)]


class GpratError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


_lib = None


def lib():
    """Load libgprat_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(f"{LIBRARY_PATH} is missing: build the backend first (__graft_entry__.build())")
        l = C.CDLL(LIBRARY_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _lib = l
    return _lib


def _check(status: int):
    if status != OK:
        raise GpratError(status, lib().gprat_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _arr(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().gprat_device_count(C.byref(n)))
    return n.value


class Context:
    """One GPU's library context (streams + caching device pool)."""

    def __init__(self, device: int = 0, n_streams: int = 1):
        h = _VP()
        _check(lib().gprat_context_create(device, n_streams, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().gprat_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference-shaped entry points ------------------------------------
    def cholesky(self, train_in, theta, n_tiles, n_tile_size, n_reg):
        x = _arr(train_in)
        t = _arr(theta)
        out = np.empty(n_tiles * (n_tiles + 1) // 2 * n_tile_size * n_tile_size)
        _check(lib().gprat_cholesky(self.handle, _dptr(x), x.size, _dptr(t), n_tiles, n_tile_size, n_reg, _dptr(out)))
        return out.reshape(-1, n_tile_size, n_tile_size)

    def predict(self, train_in, train_out, test_in, theta, n_tiles, n_tile_size, m_tiles, m_tile_size, n_reg,
                variance=False, full_cov=False, out=None):
        """Fit and predict in one call. Inputs may be host arrays or
        DeviceArrays; with `out` = (mean, var) DeviceArrays the results stay
        on the device (and those are returned)."""
        (px, nx, _kx), (py, ny, _ky), (pz, nz, _kz) = _ptr_len(train_in), _ptr_len(train_out), _ptr_len(test_in)
        t = _arr(theta)
        m = m_tiles * m_tile_size
        if out is not None:
            mean, var = out
            pm, pv = mean.as_dptr(), var.as_dptr()
        else:
            mean, var = np.empty(m), np.empty(m)
            pm, pv = _dptr(mean), _dptr(var)
        args = (self.handle, px, nx, py, ny, pz, nz, _dptr(t), n_tiles, n_tile_size, m_tiles, m_tile_size, n_reg)
        if full_cov:
            _check(lib().gprat_predict_with_full_cov(*args, pm, pv))
        elif variance:
            _check(lib().gprat_predict_with_uncertainty(*args, pm, pv))
        else:
            _check(lib().gprat_predict(*args, pm))
            return mean
        return mean, var

    def predict_full_posterior(self, train_in, train_out, test_in, theta, n_tiles, n_tile_size, m_tiles, m_tile_size, n_reg):
        """(mean, the whole m x m latent posterior covariance)."""
        x, y, z, t = _arr(train_in), _arr(train_out), _arr(test_in), _arr(theta)
        m = m_tiles * m_tile_size
        mean = np.empty(m)
        cov = np.empty((m, m))
        _check(lib().gprat_predict_full_posterior(self.handle, _dptr(x), x.size, _dptr(y), y.size, _dptr(z), z.size, _dptr(t),
                                                  n_tiles, n_tile_size, m_tiles, m_tile_size, n_reg, _dptr(mean), _dptr(cov)))
        return mean, cov

    def compute_loss(self, train_in, train_out, theta, n_tiles, n_tile_size, n_reg):
        x, y, t = _arr(train_in), _arr(train_out), _arr(theta)
        out = C.c_double(0.0)
        _check(lib().gprat_compute_loss(self.handle, _dptr(x), x.size, _dptr(y), y.size, _dptr(t), n_tiles,
                                        n_tile_size, n_reg, C.byref(out)))
        return out.value

    def loss_gradient(self, train_in, train_out, theta, n_tiles, n_tile_size, n_reg, trainable=(1, 1, 1)):
        x, y, t = _arr(train_in), _arr(train_out), _arr(theta)
        tr = (C.c_int * 3)(*[int(bool(v)) for v in trainable])
        loss = C.c_double(0.0)
        grad = np.zeros(3)
        _check(lib().gprat_loss_gradient(self.handle, _dptr(x), x.size, _dptr(y), y.size, _dptr(t), n_tiles,
                                         n_tile_size, n_reg, tr, C.byref(loss), _dptr(grad)))
        return loss.value, grad

    # -- resident models -----------------------------------------------------
    def fit(self, train_in, train_out, n, n_reg, theta, precision=PRECISION_FP64):
        return Model(self, train_in, train_out, n, n_reg, theta, precision)

    # -- measurement -----------------------------------------------------------
    def set_profiling(self, enable: bool):
        _check(lib().gprat_context_set_profiling(self.handle, int(enable)))

    def kernel_stats(self, family: str):
        n = C.c_long(0)
        ms = C.c_double(0.0)
        _check(lib().gprat_context_kernel_stats(self.handle, family.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def synchronize(self):
        _check(lib().gprat_context_synchronize(self.handle))

    def calibrate(self, copy_bytes: int = 0):
        """Measured FP64 / FP32 matrix-core TFLOP/s and copy GB/s of this device."""
        f64, f32, bw = C.c_double(), C.c_double(), C.c_double()
        _check(lib().gprat_device_calibrate(self.handle, copy_bytes, C.byref(f64), C.byref(f32), C.byref(bw)))
        return {"fp64_tflops": f64.value, "fp32_tflops": f32.value, "copy_gbs": bw.value}

    # -- device buffers ------------------------------------------------------
    def device_array(self, host: np.ndarray) -> "DeviceArray":
        """Copy a float64 array into a new device buffer."""
        return DeviceArray(self, _arr(host))

    def device_empty(self, count: int) -> "DeviceArray":
        return DeviceArray(self, None, count)


class DeviceArray:
    """A float64 device buffer owned by a context (pointer usable in any entry point)."""

    def __init__(self, ctx: Context, host, count: int = 0):
        self.ctx = ctx
        self.size = host.size if host is not None else count
        p = _VP()
        _check(lib().gprat_device_alloc(ctx.handle, max(self.size, 1) * 8, C.byref(p)))
        self.ptr = p
        if host is not None and self.size:
            _check(lib().gprat_memcpy(ctx.handle, self.ptr, host.ctypes.data_as(_VP), self.size * 8))

    def as_dptr(self):
        return C.cast(self.ptr, _D)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.size)
        if self.size:
            _check(lib().gprat_memcpy(self.ctx.handle, out.ctypes.data_as(_VP), self.ptr, self.size * 8))
        return out

    def free(self):
        if self.ptr:
            lib().gprat_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr_len(a):
    """(double*, length) for a host array or a DeviceArray."""
    if isinstance(a, DeviceArray):
        return a.as_dptr(), a.size, a
    h = _arr(a)
    return _dptr(h), h.size, h


class Model:
    """A fitted GP whose factor stays in HBM until release(). Inputs may be
    host arrays or DeviceArrays."""

    def __init__(self, ctx: Context, train_in, train_out, n, n_reg, theta, precision=PRECISION_FP64):
        (xp, xl, _xk), (yp, yl, _yk) = _ptr_len(train_in), _ptr_len(train_out)
        t = _arr(theta)
        h = _VP()
        _check(lib().gprat_model_fit(ctx.handle, xp, xl, yp, yl, n, n_reg, _dptr(t), precision, C.byref(h)))
        self.ctx = ctx
        self.handle = h
        self.n = n
        self.n_reg = n_reg

    def predict(self, test_in, m, variance=True, out=None):
        """Mean (and variance). `out=(mean, var)` may hold DeviceArrays to keep
        results in HBM; then nothing is returned."""
        zp, zl, _zk = _ptr_len(test_in)
        if out is not None:
            dm, dv = out
            _check(lib().gprat_model_predict(self.handle, zp, zl, m, dm.as_dptr(),
                                             dv.as_dptr() if (variance and dv is not None) else None))
            return None
        mean = np.empty(m)
        var = np.empty(m) if variance else None
        _check(lib().gprat_model_predict(self.handle, zp, zl, m, _dptr(mean), _dptr(var) if variance else None))
        return (mean, var) if variance else mean

    def predict_full_cov(self, test_in, m):
        """(mean, the whole m x m latent posterior covariance)."""
        zp, zl, _zk = _ptr_len(test_in)
        mean = np.empty(m)
        cov = np.empty((m, m))
        _check(lib().gprat_model_predict_full_cov(self.handle, zp, zl, m, _dptr(mean), _dptr(cov)))
        return mean, cov

    def loss(self):
        out = C.c_double(0.0)
        _check(lib().gprat_model_loss(self.handle, C.byref(out)))
        return out.value

    def gradient(self, trainable=(1, 1, 1)):
        """(loss, softplus-space gradient) from the model's own factor."""
        tr = (C.c_int * 3)(*[int(bool(v)) for v in trainable])
        loss = C.c_double(0.0)
        grad = np.zeros(3)
        _check(lib().gprat_model_gradient(self.handle, tr, C.byref(loss), _dptr(grad)))
        return loss.value, grad

    def refinement(self):
        """(iterations, last relative correction) of the alpha refinement."""
        it = C.c_int(0)
        last = C.c_double(0.0)
        _check(lib().gprat_model_refinement(self.handle, C.byref(it), C.byref(last)))
        return it.value, last.value

    def predict_mean_noalloc(self, test_in, m):
        """Mean at m test points through the library's fixed device scratch
        (no device allocation)."""
        z = _arr(test_in)
        mean = np.empty(m)
        _check(lib().gprat_model_predict_mean_noalloc(self.handle, _dptr(z), z.size, m, _dptr(mean)))
        return mean

    def set_variance_refinement(self, enabled: bool):
        """Mixed-precision models: FP64 refinement of the latent variance
        (on by default); off returns the FP32 solve's unrefined variance."""
        _check(lib().gprat_model_set_variance_refinement(self.handle, 1 if enabled else 0))

    def variance_correction(self):
        """Largest FP64 correction |q1 - q0| / v of the last refined variance."""
        v = C.c_double(0.0)
        _check(lib().gprat_model_variance_correction(self.handle, C.byref(v)))
        return v.value

    def nbytes(self):
        b = _SZ(0)
        _check(lib().gprat_model_bytes(self.handle, C.byref(b)))
        return b.value

    def release(self):
        if self.handle:
            lib().gprat_model_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Batch:
    """Many independent GPs fitted by one persistent launch; factors stay in HBM.
    `problems` is a list of (train_in, train_out, n, n_reg, theta)."""

    def __init__(self, ctx: Context, problems, precision=PRECISION_FP64):
        self.ctx = ctx
        self._keep = []
        descs = (GpDesc * len(problems))()
        for i, (x, y, n, r, theta) in enumerate(problems):
            xa, ya = _arr(x), _arr(y)
            self._keep += [xa, ya]
            descs[i].train_in = _dptr(xa)
            descs[i].train_in_len = xa.size
            descs[i].train_out = _dptr(ya)
            descs[i].train_out_len = ya.size
            descs[i].n = n
            descs[i].n_reg = r
            for k in range(3):
                descs[i].theta[k] = float(theta[k])
        h = _VP()
        _check(lib().gprat_batch_fit(ctx.handle, len(problems), C.cast(descs, C.c_void_p), precision, C.byref(h)))
        self.handle = h
        self.count = len(problems)
        self._keep = []

    def losses(self):
        out = np.empty(self.count)
        _check(lib().gprat_batch_losses(self.handle, _dptr(out)))
        return out

    def refinement(self):
        """Refinement steps per GP (zeros for an FP64 batch)."""
        out = np.zeros(self.count, dtype=np.int32)
        _check(lib().gprat_batch_refinement(self.handle, out.ctypes.data_as(_I)))
        return out

    def predict(self, index, test_in, m, variance=True):
        z = _arr(test_in)
        mean = np.empty(m)
        var = np.empty(m) if variance else None
        _check(lib().gprat_batch_predict(self.handle, index, _dptr(z), z.size, m, _dptr(mean),
                                         _dptr(var) if variance else None))
        return (mean, var) if variance else mean

    def release(self):
        if self.handle:
            lib().gprat_batch_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching on one GPU (gprat_server_*).

    submit_fit / submit_predict return tickets at once; wait(ticket) returns a
    dict with the request's status, timings and, for predicts, the mean and
    variance arrays. A fit's ticket is also its model id."""

    def __init__(self, device=0, slots=4, hbm_budget=0, precision=PRECISION_MIXED, batch_max_n=2048, batch_max=256):
        cfg = ServerConfig(device, slots, hbm_budget, precision, batch_max_n, batch_max)
        h = _VP()
        _check(lib().gprat_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h
        self._outputs = {}

    def submit_fit(self, train_in, train_out, n, n_reg, theta) -> int:
        x, y = _arr(train_in), _arr(train_out)
        d = GpDesc()
        d.train_in, d.train_in_len = _dptr(x), x.size
        d.train_out, d.train_out_len = _dptr(y), y.size
        d.n, d.n_reg = n, n_reg
        for k in range(3):
            d.theta[k] = float(theta[k])
        t = _LL(0)
        _check(lib().gprat_server_submit_fit(self.handle, C.cast(C.pointer(d), C.c_void_p), C.byref(t)))
        return t.value

    def submit_predict(self, model: int, test_in, m: int, variance=True) -> int:
        z = _arr(test_in)
        mean = np.empty(m)
        var = np.empty(m) if variance else None
        t = _LL(0)
        _check(lib().gprat_server_submit_predict(self.handle, model, _dptr(z), z.size, m, _dptr(mean),
                                                 _dptr(var) if variance else None, C.byref(t)))
        self._outputs[t.value] = (mean, var)
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """The finished request as a dict, or None if it is still running after timeout_ms."""
        r = RequestResult()
        st = lib().gprat_server_wait(self.handle, ticket, float(timeout_ms), C.byref(r))
        if st == ERR_PENDING:
            return None
        _check(st)
        out = {"status": r.status, "kind": "fit" if r.kind == 0 else "predict", "model": r.model, "loss": r.loss,
               "queue_ms": r.queue_ms, "service_ms": r.service_ms, "refit": bool(r.refit), "batched": r.batched,
               "error": r.error.decode()}
        if ticket in self._outputs:
            out["mean"], out["var"] = self._outputs.pop(ticket)
        return out

    def release_model(self, model: int):
        _check(lib().gprat_server_release_model(self.handle, model))

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().gprat_server_get_stats(self.handle, C.byref(s)))
        return {name: getattr(s, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().gprat_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
