// Host-side engine of the backend: one Context per GPU (streams, caching
// device pool, profiling), resident Models (padded factor + alpha), and the
// blocked algorithms that schedule the kernels.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "panel_plan.hpp"

namespace gpamd
{

// Caching device allocator: blocks are reused by size class so repeated
// calls of the same shape never reach hipMalloc.
class DevicePool
{
  public:
    explicit DevicePool(int device) : device_(device) { }
    ~DevicePool();
    // When a device allocation fails, this is called (without the pool's
    // lock) before the pool frees its own cache and tries once more: an
    // owner of several pools can free the others' cached blocks too.
    std::function<void()> on_out_of_memory;
    void *get(std::size_t bytes);
    void put(void *p);
    void release_cached();
    // Free cached blocks until at most `keep` bytes stay cached.
    void trim(std::size_t keep);
    std::size_t cached_bytes() const
    {
        std::lock_guard<std::mutex> g(mu_);
        return cached_;
    }
    // Pools of one owner (the server's slots) may take each other's cached
    // blocks when their own cache has none that fits, instead of allocating
    // (and, when the device is full, freeing every cache, which waits for
    // the whole device once per freed block). Each pool's context must run
    // on the one stream `stream`: a block returned to a sharing pool carries
    // an event recorded there, and a sibling that takes it waits for that
    // event on the host first, so every use of the block in its former
    // owner's stream order has completed before the new owner touches it.
    void share_with(const std::vector<DevicePool *> &siblings, hipStream_t stream);
    // Blocks taken from siblings' caches so far.
    long long borrowed() const
    {
        std::lock_guard<std::mutex> g(mu_);
        return borrowed_;
    }

  private:
    struct Cached
    {
        void *p;
        hipEvent_t returned;  // recorded on the owner's stream at put (sharing pools), else null
    };
    // Removes a cached block of size class in [want, want + 25%] (own lock).
    bool take(std::size_t want, Cached &out, std::size_t &size);

    mutable std::mutex mu_;
    int device_;
    std::multimap<std::size_t, Cached> free_;
    std::map<void *, std::size_t> live_;
    std::size_t cached_ = 0;
    std::vector<DevicePool *> siblings_;
    hipStream_t stream_ = nullptr;
    std::vector<hipEvent_t> spare_;  // events of blocks reused by this pool
    long long borrowed_ = 0;
};

struct Context;
struct BatchGraph;  // batch.cpp

// RAII handle on a pooled device block.
template <class T>
class DevBuf
{
  public:
    DevBuf() = default;
    DevBuf(Context &c, std::size_t count);
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : pool_(o.pool_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        reset();
        pool_ = o.pool_;
        p_ = o.p_;
        n_ = o.n_;
        o.p_ = nullptr;
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset();
    T *get() const { return p_; }
    std::size_t size() const { return n_; }
    std::size_t bytes() const { return n_ * sizeof(T); }

  private:
    DevicePool *pool_ = nullptr;
    T *p_ = nullptr;
    std::size_t n_ = 0;
};

struct KernelStat
{
    long launches = 0;
    double ms = 0.0;
};

struct Context
{
    Context(int device, int n_streams);
    ~Context();

    int device;
    int n_streams;
    int cus = 0;  // compute units of the device
    hipStream_t main = nullptr;   // solves, predictions, everything outside the Cholesky
    hipStream_t panel = nullptr;  // Cholesky diagonal blocks (look-ahead); == main with one stream
    hipStream_t trail = nullptr;  // Cholesky trailing updates and panel rows (== main)
    DevicePool pool;
    std::recursive_mutex mu;

    DevBuf<int> info;         // smallest non-positive pivot column
    DevBuf<double> scalars;   // reduction results

    // profiling
    bool profiling = false;
    std::map<std::string, KernelStat> stats;
    struct Pending
    {
        std::string family;
        hipEvent_t begin, end;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> spare_events;
    hipEvent_t event();
    void recycle(hipEvent_t e);
    // Brackets one launch (or launch sequence) of a kernel family.
    void prof_begin(hipStream_t s, const char *family, Pending &p);
    void prof_end(hipStream_t s, Pending &p);
    void resolve_profiling();  // after a synchronize

    // device copies of the diagonal-block task schedules, by shape
    struct PanelPlanDev
    {
        DevBuf<int> tasks, wg_off;
        std::vector<PanelLaunch> launches;
        long tmp_doubles = 0;
    };
    std::map<long, PanelPlanDev> panel_plans;
    const PanelPlanDev &panel_plan(int nb, int chain_wgs, int tail_wgs);
    const PanelPlanDev &head_plan(int nbh, int nb, int wgs);
    const PanelPlanDev &cached_plan(long key, const std::function<PanelPlan()> &build);

    // captured level-path launch sequences of the batched fit, by batch
    // composition (batch.cpp), and how often each composition was seen
    std::map<std::string, std::shared_ptr<BatchGraph>> batch_graphs;
    std::map<std::string, int> batch_seen;
    unsigned long long batch_clock = 0;

    void make_current() const;
    void sync();
};

// Brackets the launches of one kernel family with HIP events on their
// stream when the context profiles.
class Timed
{
  public:
    Timed(Context &c, hipStream_t s, const char *family) : c_(c), s_(s), on_(c.profiling)
    {
        if (on_)
            c_.prof_begin(s_, family, p_);
    }
    ~Timed()
    {
        if (on_)
            c_.prof_end(s_, p_);
    }

  private:
    Context &c_;
    hipStream_t s_;
    bool on_;
    Context::Pending p_;
};


// Resident fitted model.
struct Model
{
    Context *ctx = nullptr;
    int n = 0, n_pad = 0, r = 0;
    long lda = 0;
    SEK sek{};
    int precision = GPRAT_PRECISION_FP64;
    int refine_iterations = 0;   // mixed precision: refinement steps taken for alpha
    double refine_last = 0.0;    // last max|delta| / max|alpha|
    // mixed precision: the latent variance is refined in FP64 (the strict
    // upper triangle of L then holds the covariance, written after the
    // factorisation); variance_correction is the last prediction's
    // max |q1 - q0| / v, the FP64 correction of the FP32 solves' result
    bool refine_variance = true;
    double variance_correction = 0.0;
    DevBuf<double> x;      // training series (n + r - 1)
    DevBuf<double> y;      // targets, zero padded to n_pad
    DevBuf<double> L;      // n_pad x lda, lower triangle holds the factor
    DevBuf<double> Winv;   // inverses of the 64x64 diagonal blocks of L (n_pad/64 x 64 x 64)
    DevBuf<double> Wpanel; // k-major inverses of the factorisation's wblock-wide diagonal blocks (may be empty)
    int wblock = 0;
    DevBuf<double> alpha;  // K^-1 y (n_pad)
    std::size_t bytes() const { return x.bytes() + y.bytes() + L.bytes() + Winv.bytes() + Wpanel.bytes() + alpha.bytes(); }
};

// Raw view of a factored GP (resident model or one GP of a batch).
struct FactorView
{
    Context *ctx;
    const double *x;
    int n, n_pad, r;
    long lda;
    SEK sek;
    const double *L;
    const double *Winv;
    const double *alpha;
    bool mixed = false;               // variance solve in FP32 products (mixed-precision models)
    const double *Wpanel = nullptr;   // k-major wblock-wide diagonal-block inverses, if kept
    int wblock = 0;
    // mixed: refine the variance in FP64 (Kup: the covariance in the strict
    // upper triangle of an n_pad x n_pad matrix with leading dimension ldk,
    // or null to assemble one for the call); variance_correction receives
    // the prediction's max |q1 - q0| / v when given
    bool refine_variance = true;
    const double *Kup = nullptr;
    long ldk = 0;
    double *variance_correction = nullptr;
};

// Many GPs fitted by one persistent launch; factors stay in one arena.
struct Batch
{
    Context *ctx = nullptr;
    int count = 0;
    std::vector<k::BatchGP> gps;  // device addresses inside the arena
    DevBuf<double> arena;
    std::vector<double> losses;
    int precision = GPRAT_PRECISION_FP64;
    std::vector<int> refine_steps;  // mixed precision: refinement steps per GP
    std::vector<char> gp_mixed;     // per GP: fitted in mixed precision (a mixed batch may run some size classes in FP64)
    std::size_t bytes() const { return arena.bytes(); }
};

// Algorithms (engine.cpp)
long padded(long n);
long leading_dim(long n_pad);
void cholesky_fp64(Context &c, double *A, long lda, int n_pad, double *Winv);
// mixed: FP32 matrix-core trailing updates (FP64 panels and diagonal blocks).
// Wpanel (optional): receives the k-major inverses of the panels' diagonal
// blocks, panel_width(n_pad) wide, block k at Wpanel + k * width^2.
// extra: rows below the n_pad x n_pad matrix (A holds n_pad + extra rows)
// that the factorisation treats as rows of L below it: on return they hold
// B L^-T for the B they held (the forward solve of their transposes).
void cholesky(Context &c, double *A, long lda, int n_pad, double *Winv, bool mixed, double *Wpanel = nullptr, int extra = 0);
int panel_width(int n_pad);
void refine_alpha(Model &m, double *work);
void solve_spd(Context &c, const double *L, long lda, const double *Winv, int n_pad, const double *y, double *work, double *out,
               const double *Wpanel = nullptr, int wblock = 0);
void trsm_blocked(Context &c, const double *L, long lda, const double *Winv, const double *Wpanel, int wblock, int n_pad, double *B,
                  long ldb, int ncols, bool tri_rhs, double *out, long ldo);

void factor(Context &c, Model &m, const double *x, std::size_t x_len, int n, int r, const double theta[3], int precision);
void attach_targets(Model &m, const double *y, std::size_t y_len);
void fit(Context &c,
         Model &m,
         const double *x,
         std::size_t x_len,
         const double *y,
         std::size_t y_len,
         int n,
         int r,
         const double theta[3],
         int precision);
void predict(Model &m, const double *z, std::size_t z_len, int mt, double *mean, double *var);
// Fit and predict in one call (FP64): mean and latent variance at mt test
// points, with the rows of K* and the targets factored below K.
void predict_with_uncertainty_once(Context &c, const double *x, std::size_t x_len, const double *y, std::size_t y_len, int n, int r,
                                   const double theta[3], const double *z, std::size_t z_len, int mt, double *mean, double *var);
void predict_view(const FactorView &m, const double *z, std::size_t z_len, int mt, double *mean, double *var);
// Predictive mean through the fixed per-GPU scratch: no device allocation.
void predict_mean_noalloc(Model &m, const double *z, std::size_t z_len, int mt, double *mean);
// The view of a resident model (its context; no variance_correction output).
FactorView view(const Model &m);
// Mean and the whole mt x mt latent posterior covariance (row-major).
void predict_full_cov(Model &m, const double *z, std::size_t z_len, int mt, double *mean, double *cov);
void predict_full_cov_view(const FactorView &m, const double *z, std::size_t z_len, int mt, double *mean, double *cov);
// mixed: FP32 block products and FP64 refinement of alpha (each GP needs
// its `work` vector); refine_host (count ints) then gets the steps per GP.
// Routes each GP to the persistent kernel or the level-synchronous path and
// to FP64 or mixed precision (mixed only if requested and measured faster
// for its size class), per the autotuned table; mixed_out receives the
// per-GP precision decision.
std::vector<int> batch_run(Context &c, const std::vector<k::BatchGP> &gps, double *losses_host, bool mixed = false,
                           int *refine_host = nullptr, std::vector<char> *mixed_out = nullptr);
// Where the batched fit's tuning table came from (a file or the defaults).
std::string batch_tuning_source(int device);
void batch_fit(Context &c, Batch &b, int count, const gprat_gp_desc *gps, int precision);
void batch_losses(Batch &b, double *losses);
void batch_predict(Batch &b, int index, const double *z, std::size_t z_len, int mt, double *mean, double *var);
double loss(Model &m);
void cholesky_tiles(Model &m, int T, int b, double *out);
// Loss and softplus-space gradient of a fitted model from its own factor
// (FP64 or mixed precision).
void model_gradient(Model &m, const int trainable[3], double *loss_out, double grad[3]);
void loss_gradient(Context &c,
                   const double *x,
                   std::size_t x_len,
                   const double *y,
                   std::size_t y_len,
                   const double theta[3],
                   int n,
                   int r,
                   const int trainable[3],
                   double *loss_out,
                   double grad[3]);

}  // namespace gpamd
