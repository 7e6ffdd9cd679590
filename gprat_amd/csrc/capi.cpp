// extern "C" boundary of the backend. Converts exceptions into status codes
// plus a thread-local message, validates arguments, and serialises calls per
// context. The reference-shaped entry points (gprat_cholesky, gprat_predict,
// ...) fit a transient model, use it and release it back to the pool.
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>

#include "engine.hpp"
#include "server.hpp"

namespace
{

thread_local std::string g_last_error;

template <class Fn>
int guarded(Fn &&fn)
{
    try
    {
        fn();
        g_last_error.clear();
        return GPRAT_OK;
    }
    catch (const gpamd::Error &e)
    {
        g_last_error = e.what();
        return e.status;
    }
    catch (const std::bad_alloc &)
    {
        g_last_error = "host allocation failed";
        return GPRAT_ERR_OUT_OF_MEMORY;
    }
    catch (const std::exception &e)
    {
        g_last_error = e.what();
        return GPRAT_ERR_INTERNAL;
    }
    catch (...)
    {
        g_last_error = "unknown error";
        return GPRAT_ERR_INTERNAL;
    }
}

// Handles. A context is shared by its handle and by every resident model
// fitted on it, so destroying a context while models are alive only drops
// the handle's reference; the last model release frees it.
struct ContextHandle
{
    std::shared_ptr<gpamd::Context> ctx;
};

struct ModelHandle
{
    std::shared_ptr<gpamd::Context> owner;
    std::unique_ptr<gpamd::Model> model;
};

struct BatchHandle
{
    std::shared_ptr<gpamd::Context> owner;
    std::unique_ptr<gpamd::Batch> batch;
};

gpamd::Context &ctx_ref(gprat_context *c)
{
    gpamd::require(c != nullptr, "context is null");
    return *reinterpret_cast<ContextHandle *>(c)->ctx;
}

std::shared_ptr<gpamd::Context> ctx_shared(gprat_context *c)
{
    gpamd::require(c != nullptr, "context is null");
    return reinterpret_cast<ContextHandle *>(c)->ctx;
}

gpamd::Model &model_ref(gprat_model *m)
{
    gpamd::require(m != nullptr, "model is null");
    return *reinterpret_cast<ModelHandle *>(m)->model;
}

const gpamd::Model &model_ref(const gprat_model *m)
{
    gpamd::require(m != nullptr, "model is null");
    return *reinterpret_cast<const ModelHandle *>(m)->model;
}

struct Tiling
{
    int n, m;
};

// n = n_tiles * n_tile_size, m = m_tiles * m_tile_size (core/src/utils_c.cpp:8-52).
Tiling check_tiling(int n_tiles, int n_tile_size, int m_tiles, int m_tile_size, int n_reg)
{
    gpamd::require(n_tiles >= 1 && n_tile_size >= 1, "n_tiles and n_tile_size must be >= 1");
    gpamd::require(m_tiles >= 0 && m_tile_size >= 0, "m_tiles and m_tile_size must be >= 0");
    gpamd::require(n_reg >= 1, "n_reg must be >= 1");
    const long n = static_cast<long>(n_tiles) * n_tile_size;
    const long m = static_cast<long>(m_tiles) * m_tile_size;
    gpamd::require(n <= (1L << 30) && m <= (1L << 30), "problem too large");
    return { static_cast<int>(n), static_cast<int>(m) };
}

void fit_transient(gpamd::Context &c,
                   gpamd::Model &model,
                   const double *train_in,
                   size_t train_in_len,
                   const double *train_out,
                   size_t train_out_len,
                   const double theta[3],
                   int n,
                   int n_reg)
{
    gpamd::fit(c, model, train_in, train_in_len, train_out, train_out_len, n, n_reg, theta, GPRAT_PRECISION_FP64);
}

}  // namespace

struct gprat_context
{
};
struct gprat_model
{
};

extern "C" {

const char *gprat_last_error(void) { return g_last_error.c_str(); }

const char *gprat_version(void) { return "gprat_amd 0.1 (gfx950)"; }

int gprat_device_count(int *count)
{
    return guarded(
        [&]
        {
            gpamd::require(count != nullptr, "count is null");
            int n = 0;
            hipError_t e = hipGetDeviceCount(&n);
            if (e != hipSuccess)
            {
                (void)hipGetLastError();
                n = 0;
            }
            *count = n;
        });
}

int gprat_device_describe(int device, char *buf, size_t buf_len)
{
    return guarded(
        [&]
        {
            gpamd::require(buf != nullptr && buf_len > 0, "buffer is null");
            hipDeviceProp_t p;
            GP_HIP(hipGetDeviceProperties(&p, device));
            std::snprintf(buf,
                          buf_len,
                          "GPU %d: %s (%s), %d CUs, %.1f GiB HBM, %d MHz",
                          device,
                          p.name,
                          p.gcnArchName,
                          p.multiProcessorCount,
                          static_cast<double>(p.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0),
                          p.clockRate / 1000);
        });
}

int gprat_context_create(int device, int n_streams, gprat_context **out)
{
    return guarded(
        [&]
        {
            gpamd::require(out != nullptr, "out is null");
            gpamd::require(n_streams >= 1, "n_streams must be >= 1");
            int n = 0;
            if (hipGetDeviceCount(&n) != hipSuccess)
            {
                (void)hipGetLastError();
                n = 0;
            }
            if (n == 0)
                throw gpamd::Error(GPRAT_ERR_NO_DEVICE, "no HIP device available");
            if (device < 0 || device >= n)
                throw gpamd::Error(GPRAT_ERR_NO_DEVICE, "Requested GPU device is not available.");
            auto h = std::make_unique<ContextHandle>();
            h->ctx = std::make_shared<gpamd::Context>(device, std::max(2, n_streams));
            *out = reinterpret_cast<gprat_context *>(h.release());
        });
}

int gprat_context_destroy(gprat_context *ctx)
{
    return guarded([&] { delete reinterpret_cast<ContextHandle *>(ctx); });
}

int gprat_device_alloc(gprat_context *ctx, size_t bytes, void **ptr)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            gpamd::require(ptr != nullptr, "ptr is null");
            std::lock_guard<std::recursive_mutex> g(c.mu);
            c.make_current();
            void *p = nullptr;
            GP_HIP(hipMalloc(&p, bytes ? bytes : 1));
            *ptr = p;
        });
}

int gprat_device_free(gprat_context *ctx, void *ptr)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            std::lock_guard<std::recursive_mutex> g(c.mu);
            c.make_current();
            c.sync();
            GP_HIP(hipFree(ptr));
        });
}

int gprat_memcpy(gprat_context *ctx, void *dst, const void *src, size_t bytes)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            std::lock_guard<std::recursive_mutex> g(c.mu);
            c.make_current();
            GP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c.main));
            c.sync();
        });
}

int gprat_context_synchronize(gprat_context *ctx)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            std::lock_guard<std::recursive_mutex> g(c.mu);
            c.make_current();
            c.sync();
        });
}

int gprat_cholesky(gprat_context *ctx,
                   const double *train_in,
                   size_t train_in_len,
                   const double theta[3],
                   int n_tiles,
                   int n_tile_size,
                   int n_reg,
                   double *tiles_out)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            const Tiling t = check_tiling(n_tiles, n_tile_size, 0, 0, n_reg);
            gpamd::require(tiles_out != nullptr, "tiles_out is null");
            std::lock_guard<std::recursive_mutex> g(c.mu);
            gpamd::Model model;
            gpamd::factor(c, model, train_in, train_in_len, t.n, n_reg, theta, GPRAT_PRECISION_FP64);
            gpamd::cholesky_tiles(model, n_tiles, n_tile_size, tiles_out);
        });
}

static int predict_common(gprat_context *ctx,
                          const double *train_in,
                          size_t train_in_len,
                          const double *train_out,
                          size_t train_out_len,
                          const double *test_in,
                          size_t test_in_len,
                          const double theta[3],
                          int n_tiles,
                          int n_tile_size,
                          int m_tiles,
                          int m_tile_size,
                          int n_reg,
                          double *mean_out,
                          double *var_out,
                          bool want_var)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            const Tiling t = check_tiling(n_tiles, n_tile_size, m_tiles, m_tile_size, n_reg);
            if (t.m > 0)
            {
                gpamd::require(mean_out != nullptr, "mean_out is null");
                gpamd::require(!want_var || var_out != nullptr, "var_out is null");
            }
            std::lock_guard<std::recursive_mutex> g(c.mu);
            if (want_var && t.m > 0)
            {
                // the test rows and the targets ride along in the factorisation
                gpamd::predict_with_uncertainty_once(c, train_in, train_in_len, train_out, train_out_len, t.n, n_reg, theta, test_in,
                                                     test_in_len, t.m, mean_out, var_out);
                return;
            }
            gpamd::Model model;
            fit_transient(c, model, train_in, train_in_len, train_out, train_out_len, theta, t.n, n_reg);
            gpamd::predict(model, test_in, test_in_len, t.m, mean_out, want_var ? var_out : nullptr);
        });
}

int gprat_predict(gprat_context *ctx,
                  const double *train_in,
                  size_t train_in_len,
                  const double *train_out,
                  size_t train_out_len,
                  const double *test_in,
                  size_t test_in_len,
                  const double theta[3],
                  int n_tiles,
                  int n_tile_size,
                  int m_tiles,
                  int m_tile_size,
                  int n_reg,
                  double *mean_out)
{
    return predict_common(ctx, train_in, train_in_len, train_out, train_out_len, test_in, test_in_len, theta, n_tiles,
                          n_tile_size, m_tiles, m_tile_size, n_reg, mean_out, nullptr, false);
}

int gprat_predict_with_uncertainty(gprat_context *ctx,
                                   const double *train_in,
                                   size_t train_in_len,
                                   const double *train_out,
                                   size_t train_out_len,
                                   const double *test_in,
                                   size_t test_in_len,
                                   const double theta[3],
                                   int n_tiles,
                                   int n_tile_size,
                                   int m_tiles,
                                   int m_tile_size,
                                   int n_reg,
                                   double *mean_out,
                                   double *var_out)
{
    return predict_common(ctx, train_in, train_in_len, train_out, train_out_len, test_in, test_in_len, theta, n_tiles,
                          n_tile_size, m_tiles, m_tile_size, n_reg, mean_out, var_out, true);
}

int gprat_predict_with_full_cov(gprat_context *ctx,
                                const double *train_in,
                                size_t train_in_len,
                                const double *train_out,
                                size_t train_out_len,
                                const double *test_in,
                                size_t test_in_len,
                                const double theta[3],
                                int n_tiles,
                                int n_tile_size,
                                int m_tiles,
                                int m_tile_size,
                                int n_reg,
                                double *mean_out,
                                double *var_out)
{
    // Only diag(Sigma) leaves the reference's full-covariance path
    // (core/src/cpu/gp_functions.cpp:674-695); it equals the latent variance.
    return predict_common(ctx, train_in, train_in_len, train_out, train_out_len, test_in, test_in_len, theta, n_tiles,
                          n_tile_size, m_tiles, m_tile_size, n_reg, mean_out, var_out, true);
}

int gprat_predict_full_posterior(gprat_context *ctx,
                                 const double *train_in,
                                 size_t train_in_len,
                                 const double *train_out,
                                 size_t train_out_len,
                                 const double *test_in,
                                 size_t test_in_len,
                                 const double theta[3],
                                 int n_tiles,
                                 int n_tile_size,
                                 int m_tiles,
                                 int m_tile_size,
                                 int n_reg,
                                 double *mean_out,
                                 double *cov_out)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            const Tiling t = check_tiling(n_tiles, n_tile_size, m_tiles, m_tile_size, n_reg);
            if (t.m > 0)
                gpamd::require(mean_out != nullptr && cov_out != nullptr, "mean_out or cov_out is null");
            std::lock_guard<std::recursive_mutex> g(c.mu);
            gpamd::Model model;
            fit_transient(c, model, train_in, train_in_len, train_out, train_out_len, theta, t.n, n_reg);
            gpamd::predict_full_cov(model, test_in, test_in_len, t.m, mean_out, cov_out);
        });
}

int gprat_compute_loss(gprat_context *ctx,
                       const double *train_in,
                       size_t train_in_len,
                       const double *train_out,
                       size_t train_out_len,
                       const double theta[3],
                       int n_tiles,
                       int n_tile_size,
                       int n_reg,
                       double *loss_out)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            const Tiling t = check_tiling(n_tiles, n_tile_size, 0, 0, n_reg);
            gpamd::require(loss_out != nullptr, "loss_out is null");
            std::lock_guard<std::recursive_mutex> g(c.mu);
            gpamd::Model model;
            fit_transient(c, model, train_in, train_in_len, train_out, train_out_len, theta, t.n, n_reg);
            *loss_out = gpamd::loss(model);
        });
}

int gprat_loss_gradient(gprat_context *ctx,
                        const double *train_in,
                        size_t train_in_len,
                        const double *train_out,
                        size_t train_out_len,
                        const double theta[3],
                        int n_tiles,
                        int n_tile_size,
                        int n_reg,
                        const int trainable[3],
                        double *loss_out,
                        double grad_out[3])
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            const Tiling t = check_tiling(n_tiles, n_tile_size, 0, 0, n_reg);
            gpamd::require(loss_out != nullptr && grad_out != nullptr && trainable != nullptr, "null output");
            std::lock_guard<std::recursive_mutex> g(c.mu);
            gpamd::loss_gradient(c, train_in, train_in_len, train_out, train_out_len, theta, t.n, n_reg, trainable,
                                 loss_out, grad_out);
        });
}

int gprat_model_fit(gprat_context *ctx,
                    const double *train_in,
                    size_t train_in_len,
                    const double *train_out,
                    size_t train_out_len,
                    int n,
                    int n_reg,
                    const double theta[3],
                    int precision,
                    gprat_model **out)
{
    return guarded(
        [&]
        {
            auto owner = ctx_shared(ctx);
            gpamd::require(out != nullptr, "out is null");
            std::lock_guard<std::recursive_mutex> g(owner->mu);
            auto h = std::make_unique<ModelHandle>();
            h->owner = owner;
            h->model = std::make_unique<gpamd::Model>();
            gpamd::fit(*owner, *h->model, train_in, train_in_len, train_out, train_out_len, n, n_reg, theta, precision);
            *out = reinterpret_cast<gprat_model *>(h.release());
        });
}

int gprat_model_predict(gprat_model *model, const double *test_in, size_t test_in_len, int m, double *mean_out, double *var_out)
{
    return guarded(
        [&]
        {
            auto &md = model_ref(model);
            std::lock_guard<std::recursive_mutex> g(md.ctx->mu);
            gpamd::predict(md, test_in, test_in_len, m, mean_out, var_out);
        });
}

int gprat_model_predict_mean_noalloc(gprat_model *model, const double *test_in, size_t test_in_len, int m, double *mean_out)
{
    return guarded(
        [&]
        {
            auto &md = model_ref(model);
            std::lock_guard<std::recursive_mutex> g(md.ctx->mu);
            gpamd::predict_mean_noalloc(md, test_in, test_in_len, m, mean_out);
        });
}

int gprat_model_predict_full_cov(gprat_model *model, const double *test_in, size_t test_in_len, int m, double *mean_out, double *cov_out)
{
    return guarded(
        [&]
        {
            auto &md = model_ref(model);
            std::lock_guard<std::recursive_mutex> g(md.ctx->mu);
            gpamd::predict_full_cov(md, test_in, test_in_len, m, mean_out, cov_out);
        });
}

int gprat_model_loss(gprat_model *model, double *loss_out)
{
    return guarded(
        [&]
        {
            gpamd::require(loss_out != nullptr, "loss_out is null");
            auto &md = model_ref(model);
            std::lock_guard<std::recursive_mutex> g(md.ctx->mu);
            *loss_out = gpamd::loss(md);
        });
}

int gprat_model_gradient(gprat_model *model, const int trainable[3], double *loss_out, double grad_out[3])
{
    return guarded(
        [&]
        {
            gpamd::require(trainable != nullptr && loss_out != nullptr && grad_out != nullptr, "model_gradient: null argument");
            auto &md = model_ref(model);
            std::lock_guard<std::recursive_mutex> g(md.ctx->mu);
            gpamd::model_gradient(md, trainable, loss_out, grad_out);
        });
}

int gprat_model_refinement(const gprat_model *model, int *iterations, double *last_correction)
{
    return guarded(
        [&]
        {
            gpamd::require(iterations != nullptr && last_correction != nullptr, "null argument");
            const auto &md = model_ref(model);
            *iterations = md.refine_iterations;
            *last_correction = md.refine_last;
        });
}

int gprat_model_set_variance_refinement(gprat_model *model, int enabled)
{
    return guarded(
        [&]
        {
            auto &md = model_ref(model);
            std::lock_guard<std::recursive_mutex> g(md.ctx->mu);
            md.refine_variance = enabled != 0;
        });
}

int gprat_model_variance_correction(const gprat_model *model, double *correction)
{
    return guarded(
        [&]
        {
            gpamd::require(correction != nullptr, "null argument");
            *correction = model_ref(model).variance_correction;
        });
}

int gprat_model_bytes(const gprat_model *model, size_t *bytes)
{
    return guarded(
        [&]
        {
            gpamd::require(bytes != nullptr, "bytes is null");
            *bytes = model_ref(model).bytes();
        });
}

int gprat_model_release(gprat_model *model)
{
    return guarded(
        [&]
        {
            if (!model)
                return;
            std::unique_ptr<ModelHandle> h(reinterpret_cast<ModelHandle *>(model));
            std::shared_ptr<gpamd::Context> keep = h->owner;
            {
                std::lock_guard<std::recursive_mutex> g(keep->mu);
                keep->make_current();
                h->model.reset();
            }
            h.reset();
            keep.reset();  // frees the context if its handle is already gone
        });
}

int gprat_batch_fit(gprat_context *ctx, int count, const gprat_gp_desc *gps, int precision, gprat_batch **out)
{
    return guarded(
        [&]
        {
            auto owner = ctx_shared(ctx);
            gpamd::require(out != nullptr, "out is null");
            std::lock_guard<std::recursive_mutex> g(owner->mu);
            auto h = std::make_unique<BatchHandle>();
            h->owner = owner;
            h->batch = std::make_unique<gpamd::Batch>();
            gpamd::batch_fit(*owner, *h->batch, count, gps, precision);
            *out = reinterpret_cast<gprat_batch *>(h.release());
        });
}

int gprat_batch_losses(gprat_batch *batch, double *losses)
{
    return guarded(
        [&]
        {
            gpamd::require(batch != nullptr && losses != nullptr, "null argument");
            auto *h = reinterpret_cast<BatchHandle *>(batch);
            std::lock_guard<std::recursive_mutex> g(h->owner->mu);
            gpamd::batch_losses(*h->batch, losses);
        });
}

int gprat_batch_refinement(const gprat_batch *batch, int *steps)
{
    return guarded(
        [&]
        {
            gpamd::require(batch != nullptr && steps != nullptr, "null argument");
            auto *h = reinterpret_cast<const BatchHandle *>(batch);
            std::lock_guard<std::recursive_mutex> g(h->owner->mu);
            std::copy(h->batch->refine_steps.begin(), h->batch->refine_steps.end(), steps);
        });
}

int gprat_batch_predict(gprat_batch *batch, int index, const double *test_in, size_t test_in_len, int m, double *mean_out,
                        double *var_out)
{
    return guarded(
        [&]
        {
            gpamd::require(batch != nullptr, "batch is null");
            auto *h = reinterpret_cast<BatchHandle *>(batch);
            std::lock_guard<std::recursive_mutex> g(h->owner->mu);
            gpamd::batch_predict(*h->batch, index, test_in, test_in_len, m, mean_out, var_out);
        });
}

int gprat_batch_release(gprat_batch *batch)
{
    return guarded(
        [&]
        {
            if (!batch)
                return;
            std::unique_ptr<BatchHandle> h(reinterpret_cast<BatchHandle *>(batch));
            std::shared_ptr<gpamd::Context> keep = h->owner;
            {
                std::lock_guard<std::recursive_mutex> g(keep->mu);
                keep->make_current();
                h->batch.reset();
            }
            h.reset();
            keep.reset();
        });
}

int gprat_context_set_profiling(gprat_context *ctx, int enable)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            std::lock_guard<std::recursive_mutex> g(c.mu);
            c.make_current();
            c.sync();
            c.profiling = enable != 0;
            c.stats.clear();
        });
}

int gprat_context_kernel_stats(gprat_context *ctx, const char *family, long *launches, double *total_ms)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            gpamd::require(family != nullptr && launches != nullptr && total_ms != nullptr, "null argument");
            std::lock_guard<std::recursive_mutex> g(c.mu);
            c.make_current();
            c.sync();
            auto it = c.stats.find(family);
            *launches = it == c.stats.end() ? 0 : it->second.launches;
            *total_ms = it == c.stats.end() ? 0.0 : it->second.ms;
        });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving

struct gprat_server
{
    std::unique_ptr<gpamd::Server> impl;
};

extern "C" int gprat_server_create(const gprat_server_config *config, gprat_server **out)
{
    return guarded(
        [&]
        {
            gpamd::require(config != nullptr && out != nullptr, "server: null argument");
            *out = nullptr;
            gpamd::ServerConfig c;
            c.device = config->device;
            c.slots = config->slots;
            c.budget = config->hbm_budget;
            c.precision = config->precision;
            c.batch_max_n = config->batch_max_n;
            c.batch_max = config->batch_max;
            auto h = std::make_unique<gprat_server>();
            h->impl = std::make_unique<gpamd::Server>(c);
            *out = h.release();
        });
}

extern "C" int gprat_server_destroy(gprat_server *server)
{
    return guarded(
        [&]
        {
            gpamd::require(server != nullptr, "server: null handle");
            delete server;
        });
}

extern "C" int gprat_server_submit_fit(gprat_server *server, const gprat_gp_desc *gp, long long *ticket)
{
    return guarded(
        [&]
        {
            gpamd::require(server != nullptr && gp != nullptr && ticket != nullptr, "server: null argument");
            *ticket = server->impl->submit_fit(*gp);
        });
}

extern "C" int gprat_server_submit_predict(gprat_server *server,
                                           long long model,
                                           const double *test_in,
                                           size_t test_in_len,
                                           int m,
                                           double *mean_out,
                                           double *var_out,
                                           long long *ticket)
{
    return guarded(
        [&]
        {
            gpamd::require(server != nullptr && ticket != nullptr, "server: null argument");
            *ticket = server->impl->submit_predict(model, test_in, test_in_len, m, mean_out, var_out);
        });
}

extern "C" int gprat_server_wait(gprat_server *server, long long ticket, double timeout_ms, gprat_request_result *result)
{
    int pending = 0;
    const int st = guarded(
        [&]
        {
            gpamd::require(server != nullptr && result != nullptr, "server: null argument");
            gpamd::RequestResult r;
            if (!server->impl->wait(ticket, timeout_ms, r))
            {
                pending = 1;
                return;
            }
            std::memset(result, 0, sizeof *result);
            result->status = r.status;
            result->kind = r.kind;
            result->model = r.model;
            result->loss = r.loss;
            result->queue_ms = r.queue_ms;
            result->service_ms = r.service_ms;
            result->refit = r.refit;
            result->batched = r.batched;
            std::snprintf(result->error, sizeof result->error, "%s", r.error.c_str());
        });
    if (st == GPRAT_OK && pending)
    {
        g_last_error = "request " + std::to_string(ticket) + " has not finished";
        return GPRAT_ERR_PENDING;
    }
    return st;
}

extern "C" int gprat_server_release_model(gprat_server *server, long long model)
{
    return guarded(
        [&]
        {
            gpamd::require(server != nullptr, "server: null handle");
            server->impl->release_model(model);
        });
}

extern "C" int gprat_server_get_stats(gprat_server *server, gprat_server_stats *out)
{
    return guarded(
        [&]
        {
            gpamd::require(server != nullptr && out != nullptr, "server: null argument");
            const gpamd::ServerStats s = server->impl->stats();
            out->submitted = s.submitted;
            out->completed = s.completed;
            out->failed = s.failed;
            out->fits = s.fits;
            out->predicts = s.predicts;
            out->batched_fits = s.batched_fits;
            out->batch_launches = s.batch_launches;
            out->evictions = s.evictions;
            out->refits = s.refits;
            out->resident_models = s.resident_models;
            out->resident_bytes = s.resident_bytes;
            out->reserved_bytes = s.reserved_bytes;
            out->peak_bytes = s.peak_bytes;
            out->budget = s.budget;
            out->cache_flushes = s.cache_flushes;
            out->borrowed_blocks = s.borrowed_blocks;
        });
}

// ---------------------------------------------------------------------------
// Calibration

extern "C" int gprat_device_calibrate(gprat_context *ctx, size_t copy_bytes, double *fp64_tflops, double *fp32_tflops, double *copy_gbs)
{
    return guarded(
        [&]
        {
            auto &c = ctx_ref(ctx);
            gpamd::require(fp64_tflops != nullptr && fp32_tflops != nullptr && copy_gbs != nullptr, "calibrate: null output");
            std::lock_guard<std::recursive_mutex> g(c.mu);
            c.make_current();
            hipDeviceProp_t prop;
            GP_HIP(hipGetDeviceProperties(&prop, c.device));
            const int wgs = 8 * prop.multiProcessorCount;
            gpamd::DevBuf<double> out(c, static_cast<std::size_t>(wgs) * 256);
            hipEvent_t a, b;
            GP_HIP(hipEventCreate(&a));
            GP_HIP(hipEventCreate(&b));
            auto time_ms = [&](auto &&launch)
            {
                launch();  // warm-up
                GP_HIP(hipEventRecord(a, c.main));
                launch();
                GP_HIP(hipEventRecord(b, c.main));
                GP_HIP(hipEventSynchronize(b));
                float ms = 0.0f;
                GP_HIP(hipEventElapsedTime(&ms, a, b));
                return static_cast<double>(ms);
            };
            long flops = 0;
            double ms = time_ms([&] { flops = gpamd::k::mfma_loop(c.main, true, out.get(), wgs, 1000); });
            *fp64_tflops = flops / (ms * 1e-3) / 1e12;
            ms = time_ms([&] { flops = gpamd::k::mfma_loop(c.main, false, out.get(), wgs, 2000); });
            *fp32_tflops = flops / (ms * 1e-3) / 1e12;
            const std::size_t bytes = copy_bytes ? (copy_bytes / 16) * 16 : (std::size_t(4) << 30);
            gpamd::DevBuf<double> src(c, bytes / 8), dst(c, bytes / 8);
            gpamd::k::fill(c.main, src.get(), static_cast<long>(bytes / 8), 1.0);
            ms = time_ms([&] { gpamd::k::copy_f64(c.main, src.get(), dst.get(), static_cast<long>(bytes / 8), 4 * prop.multiProcessorCount); });
            *copy_gbs = 2.0 * bytes / (ms * 1e-3) / 1e9;
            GP_HIP(hipEventDestroy(a));
            GP_HIP(hipEventDestroy(b));
            c.sync();
        });
}
