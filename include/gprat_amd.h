/*
 * gprat_amd — C ABI of the MI355X (gfx950) backend for GPRat's exact-GP path.
 *
 * Every entry point takes plain host pointers and sizes, returns an int
 * status (GPRAT_OK == 0) and leaves a thread-local message for
 * gprat_last_error(). Device memory is owned by the library: a context owns
 * one GPU's streams and a caching pool; a model owns a resident Cholesky
 * factor. Calls on one context are serialised by the library; distinct
 * contexts are independent.
 *
 * Conventions shared by all entry points (they restate the reference's):
 *  - theta[3] = {lengthscale, vertical_lengthscale, noise_variance}.
 *  - n = n_tiles * n_tile_size training points, m = m_tiles * m_tile_size
 *    test points (integer products, as core/src/utils_c.cpp:8-52).
 *  - `train_in` / `test_in` are 1-D series; sample i has the feature window
 *    series[i .. i+n_reg-1] (core/src/cpu/gp_algorithms.cpp:11-28), so they
 *    must hold at least n + n_reg - 1 (m + n_reg - 1) values.
 *  - `train_out` holds at least n values; y_i = train_out[i].
 *  - SE kernel v * exp(-0.5 / (l*l) * |x_i - x_j|^2), noise added on the
 *    training diagonal only; predictive variance is latent (no noise).
 *  - All arithmetic is FP64 unless a model is fitted with
 *    GPRAT_PRECISION_MIXED.
 */
#ifndef GPRAT_AMD_H
#define GPRAT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. GPRAT_ERR_NOT_SPD means the covariance was not positive
 * definite: the reference ignores this (core/src/cpu/adapter_cblas_fp64.cpp:19);
 * here it is reported, with the first failing column in the message. */
enum
{
    GPRAT_OK = 0,
    GPRAT_ERR_ARGUMENT = 1,
    GPRAT_ERR_HIP = 2,
    GPRAT_ERR_OUT_OF_MEMORY = 3,
    GPRAT_ERR_NOT_SPD = 4,
    GPRAT_ERR_NO_DEVICE = 5,
    GPRAT_ERR_INTERNAL = 6,
    GPRAT_ERR_PENDING = 7 /* gprat_server_wait: the request has not finished yet */
};

/* Factorisation precision for resident models. */
enum
{
    GPRAT_PRECISION_FP64 = 0,
    /* FP64 panel (POTRF/TRSM) and log-det, FP32-MFMA trailing updates,
     * FP64 iterative refinement of alpha (residual y - K alpha in FP64).
     * Latent variance: the solves K~^-1 K*^T on FP32 matrix cores with the
     * mixed-precision factor, then one FP64 residual (K regenerated into the
     * factor's free upper triangle at fit time) that corrects the quadratic
     * form k^T K^-1 k to second order; gprat_model_set_variance_refinement
     * switches that correction off (FP32 forward solve only). The loss's
     * log-determinant and the gradient's K^-1 come from the mixed-precision
     * factor unrefined. */
    GPRAT_PRECISION_MIXED = 1
};

typedef struct gprat_context gprat_context;
typedef struct gprat_model gprat_model;

/* Message of the last failing call on this thread ("" if none). */
const char *gprat_last_error(void);

/* Library version string. */
const char *gprat_version(void);

/* Replaces gprat::gpu_count (core/include/target.hpp:202-204,
 * core/src/target.cpp:127-136). */
int gprat_device_count(int *count);

/* One-line description of a device; replaces the per-device text of
 * gprat::print_available_gpus (core/src/target.cpp:112-125). */
int gprat_device_describe(int device, char *buf, size_t buf_len);

/* Replaces gprat::CUDA_GPU(id, n_streams) plus CUDA_GPU::create/destroy
 * (core/include/target.hpp:112-180, core/src/target.cpp:24-72). n_streams
 * >= 1; the library uses at least two streams internally for look-ahead. */
int gprat_context_create(int device, int n_streams, gprat_context **out);
int gprat_context_destroy(gprat_context *ctx);

/* Device buffers for callers without their own allocator. Every pointer
 * argument of the entry points below may be host memory or device memory of
 * the context's GPU (HIP unified addressing): with device pointers the
 * inputs never cross PCIe. */
int gprat_device_alloc(gprat_context *ctx, size_t bytes, void **ptr);
int gprat_device_free(gprat_context *ctx, void *ptr);
/* Synchronous copy; either side may be host or device memory. */
int gprat_memcpy(gprat_context *ctx, void *dst, const void *src, size_t bytes);

/* Wait for all work of the context (no-op for the synchronous calls below,
 * which return after their results are in host memory). */
int gprat_context_synchronize(gprat_context *ctx);

/* Replaces gpu::cholesky (core/include/gpu/gp_functions.cuh:189-195).
 * tiles_out receives the n_tiles*(n_tiles+1)/2 lower tiles (i, j<=i) in
 * row-major tile order, each n_tile_size^2 values row-major. Diagonal tiles
 * hold L in their lower triangle and the covariance K in their strict upper
 * triangle, as the reference's POTRF leaves them. */
int gprat_cholesky(gprat_context *ctx,
                   const double *train_in,
                   size_t train_in_len,
                   const double theta[3],
                   int n_tiles,
                   int n_tile_size,
                   int n_reg,
                   double *tiles_out);

/* Replaces gpu::predict (core/include/gpu/gp_functions.cuh:27-37).
 * mean_out has m values. */
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
                  double *mean_out);

/* Replaces gpu::predict_with_uncertainty (gp_functions.cuh:55-65).
 * mean_out and var_out have m values each. Fit and prediction in one call:
 * the test points' rows of K* and a row holding the training targets are
 * factored together with K, so they come out as K* L^-T and (L^-1 y)^T
 * (the forward solves ride in the factorisation); the mean is their
 * product and the variance v minus each test row's sum of squares. The
 * results agree with a resident model's to rounding. */
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
                                   double *var_out);

/* Replaces gpu::predict_with_full_cov (gp_functions.cuh:83-93). Like the
 * reference it returns the mean and the diagonal of the posterior
 * covariance (core/src/cpu/gp_functions.cpp:663-695), computed as
 * gprat_predict_with_uncertainty computes them. */
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
                                double *var_out);

/* Mean and the whole m x m latent posterior covariance
 * Sigma = K** - K* K^-1 K*^T (row-major; cov_out holds m * m doubles).
 * The reference forms this matrix in predict_with_full_cov
 * (core/src/cpu/gp_functions.cpp:452-696, tiled_algorithms.cpp:230-254) but
 * returns only its diagonal; same arguments as gprat_predict_with_full_cov
 * otherwise. */
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
                                 double *cov_out);

/* Replaces gpu::compute_loss (gp_functions.cuh:108-114): the normalised
 * negative log marginal likelihood
 * 0.5 * (y^T K^-1 y + sum_i log(L_ii^2) + n log(2 pi)) / n. */
int gprat_compute_loss(gprat_context *ctx,
                       const double *train_in,
                       size_t train_in_len,
                       const double *train_out,
                       size_t train_out_len,
                       const double theta[3],
                       int n_tiles,
                       int n_tile_size,
                       int n_reg,
                       double *loss_out);

/* The device half of one optimisation step, replacing the loss and
 * gradient part of cpu::optimize_step (core/src/cpu/gp_functions.cpp:1032-1269,
 * core/src/cpu/tiled_algorithms.cpp:296-426), which the reference GPU path
 * lacks (core/src/gpu/gp_functions.cu:395-423). Writes the loss at theta and,
 * for each p with trainable[p] != 0, grad_out[p] = the gradient the
 * reference feeds to Adam: d loss / d u_p in softplus space
 * (u = log(exp(theta) - 1), noise shifted by 1e-6). Entries with
 * trainable[p] == 0 are set to 0. The Adam update itself stays on the host. */
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
                        double grad_out[3]);

/* ---- Resident models (fit once, predict many) -------------------------- */

/* Assemble K, factor it and solve alpha = K^-1 y; the factor stays in HBM.
 * n training points (any n >= 1; no tiling constraint). */
int gprat_model_fit(gprat_context *ctx,
                    const double *train_in,
                    size_t train_in_len,
                    const double *train_out,
                    size_t train_out_len,
                    int n,
                    int n_reg,
                    const double theta[3],
                    int precision,
                    gprat_model **out);

/* Predictive mean (and latent variance if var_out != NULL) at m test points. */
int gprat_model_predict(gprat_model *model,
                        const double *test_in,
                        size_t test_in_len,
                        int m,
                        double *mean_out,
                        double *var_out);

/* Predictive mean at m test points, like gprat_model_predict with
 * var_out == NULL, without any device allocation: the test series, the
 * mean's partial sums and the staged mean live in a fixed 8 MiB scratch
 * array that each GPU holds for the life of the process (m up to about
 * 340000 for large models; larger m fails with GPRAT_ERR_ARGUMENT). No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation). The scratch is shared by every call on one GPU: calls
 * on models of one context are serialised, calls through different contexts
 * on the same GPU must not overlap (they would race on the scratch). */
int gprat_model_predict_mean_noalloc(gprat_model *model, const double *test_in, size_t test_in_len, int m, double *mean_out);

/* Mean and the whole m x m latent posterior covariance (row-major, m * m
 * doubles) at m test points; see gprat_predict_full_posterior. */
int gprat_model_predict_full_cov(gprat_model *model,
                                 const double *test_in,
                                 size_t test_in_len,
                                 int m,
                                 double *mean_out,
                                 double *cov_out);

/* Loss at the fitted hyperparameters. */
int gprat_model_loss(gprat_model *model, double *loss_out);

/* Loss and softplus-space gradient at the model's hyperparameters from its
 * own factor, without refitting (same meaning as gprat_loss_gradient's
 * outputs). For mixed-precision models K^-1 comes from the mixed-precision
 * factor and alpha from the refined solve. */
int gprat_model_gradient(gprat_model *model, const int trainable[3], double *loss_out, double grad_out[3]);

/* Mixed-precision models: refinement steps taken for alpha and the last
 * relative correction max|delta| / max|alpha| (0 and 0.0 for FP64 models). */
int gprat_model_refinement(const gprat_model *model, int *iterations, double *last_correction);

/* Mixed-precision models: enabled != 0 (the default) refines the latent
 * variance in FP64 as described at GPRAT_PRECISION_MIXED; 0 returns the
 * FP32 forward solve's unrefined variance (faster, FP32-level error). No
 * effect on FP64 models. No counterpart in the reference (FP64 only). */
int gprat_model_set_variance_refinement(gprat_model *model, int enabled);

/* The last variance prediction's largest FP64 correction |q1 - q0| / v
 * (0.0 for FP64 models and before any refined prediction). */
int gprat_model_variance_correction(const gprat_model *model, double *correction);

/* Device bytes held by the model. */
int gprat_model_bytes(const gprat_model *model, size_t *bytes);

int gprat_model_release(gprat_model *model);

/* ---- Batches of independent GPs ----------------------------------------- */

/* One GP of a batch (host pointers). n training points, any n >= 1. */
typedef struct gprat_gp_desc
{
    const double *train_in;
    size_t train_in_len;
    const double *train_out;
    size_t train_out_len;
    int n;
    int n_reg;
    double theta[3];
} gprat_gp_desc;

typedef struct gprat_batch gprat_batch;

/* Fit `count` independent GPs of any sizes in one call (assembly, Cholesky,
 * alpha and loss per GP); factors stay resident. Per size class, from the
 * tuning table: a persistent kernel takes small GPs whole, and
 * level-synchronous launches over the 64x64 blocks of all large GPs take
 * the rest (replayed from a captured HIP graph when a batch composition
 * repeats on the context). GPRAT_PRECISION_MIXED, for the classes where
 * the table measured it faster: the block products of each factorisation
 * on FP32 matrix cores (diagonal blocks, storage and solves in FP64) and
 * alpha refined in FP64 against K regenerated from the series. No
 * counterpart in the reference, which fits one GP per call. */
int gprat_batch_fit(gprat_context *ctx, int count, const gprat_gp_desc *gps, int precision, gprat_batch **out);
/* losses[g] = loss of GP g at its hyperparameters (count values). */
int gprat_batch_losses(gprat_batch *batch, double *losses);
/* steps[g] = refinement steps taken for GP g's alpha (count values; 0 for
 * an FP64 batch). */
int gprat_batch_refinement(const gprat_batch *batch, int *steps);
/* Mean (and latent variance if var_out != NULL) of GP `index` at m points. */
int gprat_batch_predict(gprat_batch *batch, int index, const double *test_in, size_t test_in_len, int m, double *mean_out,
                        double *var_out);
int gprat_batch_release(gprat_batch *batch);

/* ---- Serving: request loop with continuous batching ---------------------
 * No counterpart in the reference, whose API is synchronous and keeps
 * nothing on the device between calls (core/src/gprat_c.cpp:108-152,
 * core/src/gpu/gp_functions.cu:81-87). A server owns `slots` worker threads,
 * each with its own stream, so admitted requests overlap on the GPU. A
 * request is admitted when resident factors plus the working memory of all
 * admitted requests stay within `hbm_budget`; idle factors are evicted least
 * recently used first and rebuilt from a host copy of their data when a
 * predict needs them again. Fits with n <= batch_max_n that are waiting
 * together run as one persistent batched launch (FP64); larger fits use
 * `precision`. Every kernel reduces in a fixed order, so a result does not
 * depend on what else was in flight. */

typedef struct gprat_server gprat_server;

typedef struct gprat_server_config
{
    int device;
    int slots;          /* requests in flight at once, 1..16 */
    size_t hbm_budget;  /* bytes; 0 = 70% of the device's memory */
    int precision;      /* GPRAT_PRECISION_FP64 or GPRAT_PRECISION_MIXED for single-GP fits */
    int batch_max_n;    /* fits with n <= batch_max_n are batched (0: never) */
    int batch_max;      /* at most this many fits per batched launch (>= 1) */
} gprat_server_config;

typedef struct gprat_request_result
{
    int status;         /* GPRAT_OK or the request's error (message in `error`) */
    int kind;           /* 0 fit, 1 predict */
    long long model;    /* model id (a fit's own ticket) */
    double loss;        /* fit: loss at theta */
    double queue_ms;    /* submit -> admitted */
    double service_ms;  /* admitted -> finished */
    int refit;          /* predict: the evicted factor was rebuilt first */
    int batched;        /* fit: GPs in its batched launch (0: single-GP path) */
    char error[256];
} gprat_request_result;

typedef struct gprat_server_stats
{
    long long submitted, completed, failed;
    long long fits, predicts, batched_fits, batch_launches;
    long long evictions, refits, resident_models;
    size_t resident_bytes, reserved_bytes, peak_bytes, budget;
    /* device allocations that failed and freed every slot's cached blocks;
     * cached blocks one slot reused from another slot's cache */
    long long cache_flushes, borrowed_blocks;
} gprat_server_stats;

int gprat_server_create(const gprat_server_config *config, gprat_server **out);
/* Runs every queued request to completion, then frees all models. */
int gprat_server_destroy(gprat_server *server);
/* Queue a fit (inputs copied before return). *ticket is also the model id. */
int gprat_server_submit_fit(gprat_server *server, const gprat_gp_desc *gp, long long *ticket);
/* Queue a predict against a model (its fit may still be queued). mean_out
 * (m values) and var_out (m values, or NULL) are written when the request
 * finishes and must stay valid until gprat_server_wait has returned it. */
int gprat_server_submit_predict(gprat_server *server,
                                long long model,
                                const double *test_in,
                                size_t test_in_len,
                                int m,
                                double *mean_out,
                                double *var_out,
                                long long *ticket);
/* Wait up to timeout_ms (< 0: no limit) for a request. Returns GPRAT_OK with
 * *result filled (the request's own status in result->status) and retires
 * the ticket, or GPRAT_ERR_PENDING on timeout. */
int gprat_server_wait(gprat_server *server, long long ticket, double timeout_ms, gprat_request_result *result);
/* Drop a model; its factor is freed once no admitted request uses it. */
int gprat_server_release_model(gprat_server *server, long long model);
int gprat_server_get_stats(gprat_server *server, gprat_server_stats *out);

/* ---- Measurement hooks ------------------------------------------------- */

/* When enabled, the library brackets every launch of its named kernels with
 * HIP events on the launching stream and accumulates their durations. */
int gprat_context_set_profiling(gprat_context *ctx, int enable);
/* Launch count and summed duration (ms) for one kernel family since
 * profiling was enabled. Families: "cholesky" (whole factorisation),
 * "panel" (diagonal blocks), "panel_rows", "syrk_trailing", "assemble",
 * "solve", "trsm" (variance solves, whole), "trsm_update" (their update
 * products only), "mean", "refine", "residual" (the FP64 product of the
 * variance refinement), "inverse", "gradient", "batch_fit". */
int gprat_context_kernel_stats(gprat_context *ctx, const char *family, long *launches, double *total_ms);

/* Measured ceilings of this device, to report rates against: FP64 and FP32
 * matrix-core TFLOP/s (16x16x4 MFMAs from registers on every CU) and the
 * device-to-device copy bandwidth in GB/s (read + write bytes over a
 * `copy_bytes` copy; 0 selects 4 GiB). No counterpart in the reference. */
int gprat_device_calibrate(gprat_context *ctx, size_t copy_bytes, double *fp64_tflops, double *fp32_tflops, double *copy_gbs);

#ifdef __cplusplus
}
#endif

#endif /* GPRAT_AMD_H */
