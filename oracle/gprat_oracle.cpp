// TEST INFRASTRUCTURE — the parity oracle (see gprat_oracle.hpp). Not product code.
//
// Restates GPRat's CPU GP path. The tile graph of every API call is built in
// the same order as the reference and each tile op makes the same
// CBLAS/LAPACKE call with the same arguments, so the arithmetic is the
// reference's up to the BLAS library it is linked against. HPX futures and
// `hpx::dataflow` are replaced by `Flow<T>` and `after(...)` below: a task runs
// on the pool once all the futures it consumes are ready. Each tile value is
// immutable once produced, so the schedule does not change any result.
#include "gprat_oracle.hpp"

#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <dlfcn.h>
#include <functional>
#include <memory>
#include <mutex>
#include <numbers>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace oracle
{

// ---------------------------------------------------------------------------
// Hyperparameter structs: core/src/gp_kernels.cpp:7-53
SEKParams::SEKParams(double l, double v, double n) :
    lengthscale(l), vertical_lengthscale(v), noise_variance(n), m_T(3, 0.0), w_T(3, 0.0)
{ }

void SEKParams::set_param(std::size_t index, double value)
{
    switch (index)
    {
        case 0: lengthscale = value; return;
        case 1: vertical_lengthscale = value; return;
        case 2: noise_variance = value; return;
        default: throw std::invalid_argument("Set Invalid param_idx");
    }
}

const double &SEKParams::get_param(std::size_t index) const
{
    switch (index)
    {
        case 0: return lengthscale;
        case 1: return vertical_lengthscale;
        case 2: return noise_variance;
        default: throw std::invalid_argument("Get Invalid param_idx");
    }
}

// ---------------------------------------------------------------------------
// Tiling helpers and loader: core/src/utils_c.cpp:8-81
int compute_train_tiles(int n_samples, int n_tile_size)
{
    if (n_tile_size <= 0)
        throw std::runtime_error("Error: Please specify a valid value for train_tile_size.\n");
    return n_samples / n_tile_size;
}

int compute_train_tile_size(int n_samples, int n_tiles)
{
    if (n_tiles <= 0)
        throw std::runtime_error("Error: Please specify a valid value for train_tiles.\n");
    return n_samples / n_tiles;
}

std::pair<int, int> compute_test_tiles(int n_test, int n_tiles, int n_tile_size)
{
    // Not divisible (which includes n_test < tile size): keep the tile count.
    if (n_test % n_tile_size > 0)
        return { n_tiles, n_test / n_tiles };
    return { n_test / n_tile_size, n_tile_size };
}

std::vector<double> load_data(const std::string &file_path, int n_samples, int offset)
{
    std::vector<double> data(static_cast<std::size_t>(n_samples + offset), 0.0);
    FILE *f = std::fopen(file_path.c_str(), "r");
    if (!f)
        throw std::runtime_error("Error: File not found: " + file_path);
    int scanned = 0;
    for (int i = 0; i < n_samples; i++)
        scanned += std::fscanf(f, "%lf", &data[static_cast<std::size_t>(i + offset)]);
    std::fclose(f);
    if (scanned != n_samples)
        throw std::runtime_error("Error: Data not correctly read. Expected " + std::to_string(n_samples)
                                 + " elements, but read " + std::to_string(scanned));
    return data;
}

// ---------------------------------------------------------------------------
// BLAS binding. The reference links CBLAS/LAPACKE at build time
// (core/src/cpu/adapter_cblas_fp64.cpp:3-10); here the same entry points are
// resolved at run time so the oracle builds without those headers.
namespace
{
enum : int
{
    kRowMajor = 101,
    kNoTrans = 111,
    kTrans = 112,
    kLower = 122,
    kNonUnit = 131,
    kLeft = 141,
    kRight = 142,
};

struct BlasTable
{
    std::string name;
    std::function<void(int, char, long, double *, long)> potrf2;
    std::function<void(int, int, int, int, int, long, long, double, const double *, long, double *, long)> trsm;
    std::function<void(int, int, int, long, long, double, const double *, long, double, double *, long)> syrk;
    std::function<void(int,
                       int,
                       int,
                       long,
                       long,
                       long,
                       double,
                       const double *,
                       long,
                       const double *,
                       long,
                       double,
                       double *,
                       long)>
        gemm;
    std::function<void(int, int, int, int, long, const double *, long, double *, long)> trsv;
    std::function<void(int, int, long, long, double, const double *, long, const double *, long, double, double *, long)>
        gemv;
    std::function<double(long, const double *, long, const double *, long)> ddot;
    std::function<void(long, double, const double *, long, double *, long)> daxpy;
    // single precision (core/src/cpu/adapter_cblas_fp32.cpp:14-100)
    std::function<void(int, char, long, float *, long)> spotrf2;
    std::function<void(int, int, int, int, int, long, long, float, const float *, long, float *, long)> strsm;
    std::function<void(int, int, int, long, long, float, const float *, long, float, float *, long)> ssyrk;
    std::function<void(int, int, int, long, long, long, float, const float *, long, const float *, long, float, float *, long)> sgemm;
};

template <class I>
BlasTable bind_blas(void *h, const std::string &name, const char *pre, const char *suf)
{
    auto sym = [&](const char *base) -> void *
    {
        std::string s = std::string(pre) + base + suf;
        void *p = dlsym(h, s.c_str());
        if (!p)
            throw std::runtime_error("oracle: missing BLAS symbol " + s);
        return p;
    };
    BlasTable t;
    t.name = name;
    auto potrf = reinterpret_cast<I (*)(int, char, I, double *, I)>(sym("LAPACKE_dpotrf2"));
    t.potrf2 = [potrf](int lay, char up, long n, double *a, long lda) { potrf(lay, up, I(n), a, I(lda)); };
    auto trsm = reinterpret_cast<void (*)(int, int, int, int, int, I, I, double, const double *, I, double *, I)>(
        sym("cblas_dtrsm"));
    t.trsm = [trsm](int o, int s, int u, int tr, int d, long m, long n, double al, const double *a, long lda, double *b,
                    long ldb) { trsm(o, s, u, tr, d, I(m), I(n), al, a, I(lda), b, I(ldb)); };
    auto syrk =
        reinterpret_cast<void (*)(int, int, int, I, I, double, const double *, I, double, double *, I)>(sym("cblas_dsyrk"));
    t.syrk = [syrk](int o, int u, int tr, long n, long k, double al, const double *a, long lda, double be, double *c,
                    long ldc) { syrk(o, u, tr, I(n), I(k), al, a, I(lda), be, c, I(ldc)); };
    auto gemm = reinterpret_cast<void (*)(int, int, int, I, I, I, double, const double *, I, const double *, I, double,
                                          double *, I)>(sym("cblas_dgemm"));
    t.gemm = [gemm](int o, int ta, int tb, long m, long n, long k, double al, const double *a, long lda,
                    const double *b, long ldb, double be, double *c, long ldc)
    { gemm(o, ta, tb, I(m), I(n), I(k), al, a, I(lda), b, I(ldb), be, c, I(ldc)); };
    auto trsv = reinterpret_cast<void (*)(int, int, int, int, I, const double *, I, double *, I)>(sym("cblas_dtrsv"));
    t.trsv = [trsv](int o, int u, int tr, int d, long n, const double *a, long lda, double *x, long incx)
    { trsv(o, u, tr, d, I(n), a, I(lda), x, I(incx)); };
    auto gemv = reinterpret_cast<void (*)(int, int, I, I, double, const double *, I, const double *, I, double,
                                          double *, I)>(sym("cblas_dgemv"));
    t.gemv = [gemv](int o, int tr, long m, long n, double al, const double *a, long lda, const double *x, long incx,
                    double be, double *y, long incy)
    { gemv(o, tr, I(m), I(n), al, a, I(lda), x, I(incx), be, y, I(incy)); };
    auto ddot = reinterpret_cast<double (*)(I, const double *, I, const double *, I)>(sym("cblas_ddot"));
    t.ddot = [ddot](long n, const double *x, long ix, const double *y, long iy) { return ddot(I(n), x, I(ix), y, I(iy)); };
    auto daxpy = reinterpret_cast<void (*)(I, double, const double *, I, double *, I)>(sym("cblas_daxpy"));
    t.daxpy = [daxpy](long n, double a, const double *x, long ix, double *y, long iy)
    { daxpy(I(n), a, x, I(ix), y, I(iy)); };
    auto spotrf = reinterpret_cast<I (*)(int, char, I, float *, I)>(sym("LAPACKE_spotrf2"));
    t.spotrf2 = [spotrf](int lay, char up, long n, float *a, long lda) { spotrf(lay, up, I(n), a, I(lda)); };
    auto strsm = reinterpret_cast<void (*)(int, int, int, int, int, I, I, float, const float *, I, float *, I)>(sym("cblas_strsm"));
    t.strsm = [strsm](int o, int s, int u, int tr, int d, long m, long n, float al, const float *a, long lda, float *b, long ldb)
    { strsm(o, s, u, tr, d, I(m), I(n), al, a, I(lda), b, I(ldb)); };
    auto ssyrk = reinterpret_cast<void (*)(int, int, int, I, I, float, const float *, I, float, float *, I)>(sym("cblas_ssyrk"));
    t.ssyrk = [ssyrk](int o, int u, int tr, long n, long k, float al, const float *a, long lda, float be, float *c, long ldc)
    { ssyrk(o, u, tr, I(n), I(k), al, a, I(lda), be, c, I(ldc)); };
    auto sgemm = reinterpret_cast<void (*)(int, int, int, I, I, I, float, const float *, I, const float *, I, float, float *, I)>(
        sym("cblas_sgemm"));
    t.sgemm = [sgemm](int o, int ta, int tb, long m, long n, long k, float al, const float *a, long lda, const float *b, long ldb,
                      float be, float *c, long ldc) { sgemm(o, ta, tb, I(m), I(n), I(k), al, a, I(lda), b, I(ldb), be, c, I(ldc)); };
    return t;
}

std::string numpy_openblas_path()
{
    // numpy wheels ship scipy-openblas (ILP64, prefixed symbols) in numpy.libs.
    const char *env = std::getenv("GPRAT_ORACLE_OPENBLAS");
    if (env)
        return env;
    const char *dirs[] = { "/usr/local/lib/python3.10/dist-packages/numpy.libs", "/usr/lib/python3/dist-packages/numpy.libs" };
    for (const char *d : dirs)
    {
        std::string cmd = std::string("ls ") + d + "/libscipy_openblas64_*.so 2>/dev/null";
        FILE *p = popen(cmd.c_str(), "r");
        if (!p)
            continue;
        char buf[1024] = { 0 };
        bool ok = std::fgets(buf, sizeof buf, p) != nullptr;
        pclose(p);
        if (ok)
        {
            std::string s(buf);
            while (!s.empty() && (s.back() == '\n' || s.back() == '\r'))
                s.pop_back();
            return s;
        }
    }
    return "";
}

BlasTable load_blas()
{
    const char *force = std::getenv("GPRAT_ORACLE_BLAS");
    std::string want = force ? force : "";
    if (want.empty() || want == "mkl")
    {
        const char *paths[] = { "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1" };
        for (const char *p : paths)
        {
            void *h = dlopen(p, RTLD_NOW | RTLD_GLOBAL);
            if (!h)
                continue;
            // MKL_THREADING sequential, as the reference builds it (CMakeLists.txt:53).
            if (auto set_layer = reinterpret_cast<int (*)(int)>(dlsym(h, "MKL_Set_Threading_Layer")))
                set_layer(1 /* MKL_THREADING_SEQUENTIAL */);
            return bind_blas<int>(h, "mkl", "", "");
        }
        if (want == "mkl")
            throw std::runtime_error("oracle: GPRAT_ORACLE_BLAS=mkl but libmkl_rt.so not found");
    }
    std::string path = numpy_openblas_path();
    void *h = path.empty() ? nullptr : dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h)
        throw std::runtime_error("oracle: no BLAS found (neither libmkl_rt.so nor numpy's scipy-openblas)");
    if (auto set_threads = reinterpret_cast<void (*)(int)>(dlsym(h, "scipy_openblas_set_num_threads64_")))
        set_threads(1);
    BlasTable t = bind_blas<long>(h, "openblas64", "scipy_", "64_");
    // LAPACKE symbol is scipy_LAPACKE_dpotrf264_ (the "2" belongs to the routine name).
    return t;
}

const BlasTable &blas()
{
    static BlasTable table = load_blas();
    return table;
}

// ---------------------------------------------------------------------------
// Dataflow pool standing in for the HPX runtime.
class Pool
{
  public:
    explicit Pool(unsigned n)
    {
        for (unsigned i = 0; i < n; i++)
            workers_.emplace_back([this] { loop(); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &w : workers_)
            w.join();
    }
    void submit(std::function<void()> f)
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }
    unsigned size() const { return static_cast<unsigned>(workers_.size()); }

  private:
    void loop()
    {
        for (;;)
        {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [this] { return stop_ || !q_.empty(); });
                if (q_.empty())
                    return;
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
        }
    }
    std::vector<std::thread> workers_;
    std::deque<std::function<void()>> q_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false;
};

std::mutex g_pool_mu;
std::unique_ptr<Pool> g_pool;

Pool &pool()
{
    std::lock_guard<std::mutex> g(g_pool_mu);
    if (!g_pool)
    {
        unsigned n = std::thread::hardware_concurrency();
        g_pool = std::make_unique<Pool>(n ? n : 1);
    }
    return *g_pool;
}

// A value that becomes ready once; continuations run on readiness.
template <class T>
class Flow
{
  public:
    struct State
    {
        std::mutex mu;
        std::condition_variable cv;
        bool ready = false;
        T value{};
        std::vector<std::function<void()>> waiters;
    };

    Flow() = default;
    static Flow now(T v)
    {
        Flow f;
        f.st_ = std::make_shared<State>();
        f.st_->value = std::move(v);
        f.st_->ready = true;
        return f;
    }
    static Flow pending()
    {
        Flow f;
        f.st_ = std::make_shared<State>();
        return f;
    }
    bool valid() const { return static_cast<bool>(st_); }
    const T &get() const
    {
        std::unique_lock<std::mutex> g(st_->mu);
        st_->cv.wait(g, [this] { return st_->ready; });
        return st_->value;
    }
    void when_ready(std::function<void()> f) const
    {
        {
            std::lock_guard<std::mutex> g(st_->mu);
            if (!st_->ready)
            {
                st_->waiters.push_back(std::move(f));
                return;
            }
        }
        f();
    }
    void fulfil(T v) const
    {
        std::vector<std::function<void()>> run;
        {
            std::lock_guard<std::mutex> g(st_->mu);
            st_->value = std::move(v);
            st_->ready = true;
            run.swap(st_->waiters);
        }
        st_->cv.notify_all();
        for (auto &f : run)
            f();
    }

  private:
    std::shared_ptr<State> st_;
};

// Runs `fn(deps.get()...)` on the pool once every dependency is ready.
template <class Fn, class... Ts>
auto after(Fn fn, Flow<Ts>... deps) -> Flow<std::decay_t<decltype(fn(deps.get()...))>>
{
    using R = std::decay_t<decltype(fn(deps.get()...))>;
    Flow<R> out = Flow<R>::pending();
    auto remaining = std::make_shared<std::atomic<int>>(static_cast<int>(sizeof...(Ts)) + 1);
    auto launch = [=]() { pool().submit([=]() { out.fulfil(fn(deps.get()...)); }); };
    auto arrive = [=]()
    {
        if (remaining->fetch_sub(1) == 1)
            launch();
    };
    (deps.when_ready(arrive), ...);
    arrive();
    return out;
}

// Runs `fn(values)` on the pool once every future in `deps` is ready.
template <class Fn, class T>
auto after_all(Fn fn, std::vector<Flow<T>> deps) -> Flow<std::decay_t<decltype(fn(std::vector<T>{}))>>
{
    using R = std::decay_t<decltype(fn(std::vector<T>{}))>;
    Flow<R> out = Flow<R>::pending();
    auto remaining = std::make_shared<std::atomic<int>>(static_cast<int>(deps.size()) + 1);
    auto launch = [=]()
    {
        pool().submit(
            [=]()
            {
                std::vector<T> vals;
                vals.reserve(deps.size());
                for (auto &d : deps)
                    vals.push_back(d.get());
                out.fulfil(fn(vals));
            });
    };
    auto arrive = [=]()
    {
        if (remaining->fetch_sub(1) == 1)
            launch();
    };
    for (auto &d : deps)
        d.when_ready(arrive);
    arrive();
    return out;
}

// Tiles are immutable vectors shared between consumers.
using Vec = std::vector<double>;
using Tile = std::shared_ptr<const Vec>;
using FTile = Flow<Tile>;
using Tiles = std::vector<FTile>;
using FScalar = Flow<double>;

Tile make_tile(Vec v) { return std::make_shared<const Vec>(std::move(v)); }

template <class Fn>
FTile spawn(Fn fn)
{
    return after([fn]() { return make_tile(fn()); });
}

// ---------------------------------------------------------------------------
// Tile generation: core/src/cpu/gp_algorithms.cpp:11-183

// gp_algorithms.cpp:11-28. Distance accumulated in k order, then
// v * exp(-0.5 / (l*l) * d) evaluated left to right.
double covariance(std::size_t ig, std::size_t jg, std::size_t r, const SEKParams &s, const Vec &xi, const Vec &xj)
{
    double d = 0.0;
    for (std::size_t k = 0; k < r; k++)
    {
        double t = xi[ig + k] - xj[jg + k];
        d += t * t;
    }
    return s.vertical_lengthscale * std::exp(-0.5 / (s.lengthscale * s.lengthscale) * d);
}

// gp_algorithms.cpp:30-62 (noise on the global diagonal only)
Vec tile_covariance(std::size_t row, std::size_t col, std::size_t N, std::size_t r, const SEKParams &s, const Vec &x)
{
    Vec t;
    t.reserve(N * N);
    for (std::size_t i = 0; i < N; i++)
        for (std::size_t j = 0; j < N; j++)
        {
            std::size_t ig = N * row + i, jg = N * col + j;
            double c = covariance(ig, jg, r, s, x, x);
            if (ig == jg)
                c += s.noise_variance;
            t.push_back(c);
        }
    return t;
}

// gp_algorithms.cpp:64-88
Vec tile_full_prior(std::size_t row, std::size_t col, std::size_t N, std::size_t r, const SEKParams &s, const Vec &x)
{
    Vec t;
    t.reserve(N * N);
    for (std::size_t i = 0; i < N; i++)
        for (std::size_t j = 0; j < N; j++)
            t.push_back(covariance(N * row + i, N * col + j, r, s, x, x));
    return t;
}

// gp_algorithms.cpp:90-111
Vec tile_prior_diag(std::size_t row, std::size_t col, std::size_t N, std::size_t r, const SEKParams &s, const Vec &x)
{
    Vec t;
    t.reserve(N);
    for (std::size_t i = 0; i < N; i++)
        t.push_back(covariance(N * row + i, N * col + i, r, s, x, x));
    return t;
}

// gp_algorithms.cpp:113-140
Vec tile_cross(std::size_t row,
               std::size_t col,
               std::size_t Nr,
               std::size_t Nc,
               std::size_t r,
               const SEKParams &s,
               const Vec &xr,
               const Vec &xc)
{
    Vec t;
    t.reserve(Nr * Nc);
    for (std::size_t i = 0; i < Nr; i++)
        for (std::size_t j = 0; j < Nc; j++)
            t.push_back(covariance(Nr * row + i, Nc * col + j, r, s, xr, xc));
    return t;
}

// gp_algorithms.cpp:142-157: (Nr x Nc) -> (Nc x Nr)
Vec tile_transpose(std::size_t Nr, std::size_t Nc, const Vec &a)
{
    Vec t;
    t.reserve(Nr * Nc);
    for (std::size_t j = 0; j < Nc; j++)
        for (std::size_t i = 0; i < Nr; i++)
            t.push_back(a[i * Nc + j]);
    return t;
}

// gp_algorithms.cpp:159-169
Vec tile_output(std::size_t row, std::size_t N, const Vec &y)
{
    return Vec(y.begin() + static_cast<long>(N * row), y.begin() + static_cast<long>(N * (row + 1)));
}

// gp_algorithms.cpp:171-183
Vec tile_zeros(std::size_t N) { return Vec(N, 0.0); }
Vec tile_identity(std::size_t N)
{
    Vec t(N * N, 0.0);
    for (std::size_t i = 0; i < N; i++)
        t[i * N + i] = 1.0;
    return t;
}

// ---------------------------------------------------------------------------
// Optimizer scalars and gradient tiles: core/src/cpu/gp_optimizer.cpp:12-245
}  // namespace

double to_constrained(double p, bool noise)  // gp_optimizer.cpp:12-22
{
    return noise ? std::log(1.0 + std::exp(p)) + 1e-6 : std::log(1.0 + std::exp(p));
}

double to_unconstrained(double p, bool noise)  // gp_optimizer.cpp:24-34
{
    return noise ? std::log(std::exp(p - 1e-6) - 1.0) : std::log(std::exp(p) - 1.0);
}

double compute_sigmoid(double p) { return 1.0 / (1.0 + std::exp(-p)); }  // gp_optimizer.cpp:36

namespace
{
// gp_optimizer.cpp:40-57: returns -0.5 / (l*l) * d
double distance_term(std::size_t ig, std::size_t jg, std::size_t r, const SEKParams &s, const Vec &x)
{
    double d = 0.0;
    for (std::size_t k = 0; k < r; k++)
    {
        double t = x[ig + k] - x[jg + k];
        d += t * t;
    }
    return -0.5 / (s.lengthscale * s.lengthscale) * d;
}

// gp_optimizer.cpp:59-82
Vec tile_distance(std::size_t row, std::size_t col, std::size_t N, std::size_t r, const SEKParams &s, const Vec &x)
{
    Vec t;
    t.reserve(N * N);
    for (std::size_t i = 0; i < N; i++)
        for (std::size_t j = 0; j < N; j++)
            t.push_back(distance_term(N * row + i, N * col + j, r, s, x));
    return t;
}

// gp_optimizer.cpp:84-114
Vec tile_covariance_from_distance(std::size_t row, std::size_t col, std::size_t N, const SEKParams &s, const Vec &D)
{
    Vec t;
    t.reserve(N * N);
    for (std::size_t i = 0; i < N; i++)
        for (std::size_t j = 0; j < N; j++)
        {
            double c = s.vertical_lengthscale * std::exp(D[i * N + j]);
            if (N * row + i == N * col + j)
                c += s.noise_variance;
            t.push_back(c);
        }
    return t;
}

// gp_optimizer.cpp:115-132: dK/dv with the softplus chain factor baked in
Vec tile_grad_v(std::size_t N, const SEKParams &s, const Vec &D)
{
    const double der = compute_sigmoid(to_unconstrained(s.vertical_lengthscale, false));
    Vec t;
    t.reserve(N * N);
    for (std::size_t e = 0; e < N * N; e++)
        t.push_back(std::exp(D[e]) * der);
    return t;
}

// gp_optimizer.cpp:133-150: dK/dl = -2 v / l * D * exp(D) * sigmoid(u_l)
Vec tile_grad_l(std::size_t N, const SEKParams &s, const Vec &D)
{
    const double der = compute_sigmoid(to_unconstrained(s.lengthscale, false));
    const double factor = -2.0 * s.vertical_lengthscale / s.lengthscale;
    Vec t;
    t.reserve(N * N);
    for (std::size_t e = 0; e < N * N; e++)
        t.push_back(factor * D[e] * std::exp(D[e]) * der);
    return t;
}

// gp_optimizer.cpp:154-183 (Adam, "option 2")
double first_moment(double g, double m, double b1) { return b1 * m + (1.0 - b1) * g; }
double second_moment(double g, double v, double b2) { return b2 * v + (1.0 - b2) * g * g; }
double adam_update(double u, const AdamParams &a, double m, double v, std::size_t iter)
{
    double b1t = std::pow(a.beta1, static_cast<double>(iter + 1));
    double b2t = std::pow(a.beta2, static_cast<double>(iter + 1));
    double nu = a.learning_rate * std::sqrt(1.0 - b2t) / (1.0 - b1t);
    return u - nu * m / (std::sqrt(v) + a.epsilon);
}

// ---------------------------------------------------------------------------
// Per-tile BLAS ops: core/src/cpu/adapter_cblas_fp64.cpp:14-194. Each takes
// its input tiles by value-copy semantics (a fresh output vector).

// :14-22 LAPACKE_dpotrf2 row-major lower, info ignored
Tile op_potrf(const Tile &A, int N)
{
    Vec a = *A;
    blas().potrf2(kRowMajor, 'L', N, a.data(), N);
    return make_tile(std::move(a));
}

// :24-52 trsm(L, A, N, M, trans, side)
Tile op_trsm(const Tile &L, const Tile &A, int N, int M, int trans, int side)
{
    Vec a = *A;
    blas().trsm(kRowMajor, side, kLower, trans, kNonUnit, N, M, 1.0, L->data(), N, a.data(), M);
    return make_tile(std::move(a));
}

// :54-65 syrk: A = A - B B^T (lower)
Tile op_syrk(const Tile &A, const Tile &B, int N)
{
    Vec a = *A;
    blas().syrk(kRowMajor, kLower, kNoTrans, N, N, -1.0, B->data(), N, 1.0, a.data(), N);
    return make_tile(std::move(a));
}

// :67-100 gemm(A, B, C, N, M, K, tA, tB): C(KxM) = C - op(A) op(B), lda = K, ldb = ldc = M
Tile op_gemm(const Tile &A, const Tile &B, const Tile &C, int N, int M, int K, int tA, int tB)
{
    Vec c = *C;
    blas().gemm(kRowMajor, tA, tB, K, M, N, -1.0, A->data(), K, B->data(), M, 1.0, c.data(), M);
    return make_tile(std::move(c));
}

// :104-120
Tile op_trsv(const Tile &L, const Tile &a, int N, int trans)
{
    Vec x = *a;
    blas().trsv(kRowMajor, kLower, trans, kNonUnit, N, L->data(), N, x.data(), 1);
    return make_tile(std::move(x));
}

// :122-152 gemv: b = b + alpha op(A) a, A is N x M row-major
Tile op_gemv(const Tile &A, const Tile &a, const Tile &b, int N, int M, double alpha, int trans)
{
    Vec y = *b;
    blas().gemv(kRowMajor, trans, N, M, alpha, A->data(), M, a->data(), 1, 1.0, y.data(), 1);
    return make_tile(std::move(y));
}

// :154-165 r[j] += column j of A (N x M) dotted with itself
Tile op_dot_diag_syrk(const Tile &A, const Tile &r, int N, int M)
{
    Vec out = *r;
    for (std::size_t j = 0; j < static_cast<std::size_t>(M); ++j)
        out[j] += blas().ddot(N, &(*A)[j], M, &(*A)[j], M);
    return make_tile(std::move(out));
}

// :167-178 r[i] += row i of A (N x M) dotted with column i of B (M x N)
Tile op_dot_diag_gemm(const Tile &A, const Tile &B, const Tile &r, int N, int M)
{
    Vec out = *r;
    for (std::size_t i = 0; i < static_cast<std::size_t>(N); ++i)
        out[i] += blas().ddot(M, &(*A)[i * static_cast<std::size_t>(M)], 1, &(*B)[i], N);
    return make_tile(std::move(out));
}

// :182-188 y - x
Tile op_axpy(const Tile &y, const Tile &x, int N)
{
    Vec out = *y;
    blas().daxpy(N, -1.0, x->data(), 1, out.data(), 1);
    return make_tile(std::move(out));
}

// :190-194
double op_dot(const Vec &a, const Vec &b, int N) { return blas().ddot(N, a.data(), 1, b.data(), 1); }

// ---------------------------------------------------------------------------
// FP32 tile ops: core/src/cpu/adapter_cblas_fp32.cpp:14-100, the same calls
// as the FP64 adapter on float tiles (LAPACKE_spotrf2, cblas_strsm/ssyrk/
// sgemm). The reference keeps this adapter beside the FP64 one without
// wiring it into a GP function; here it is the CPU comparator of the mixed-
// precision path. Tiles stay FP64 vectors between ops: each op rounds its
// inputs to float, calls the float BLAS and widens the result.

std::vector<float> to32(const Vec &a) { return std::vector<float>(a.begin(), a.end()); }
Vec from32(const std::vector<float> &a) { return Vec(a.begin(), a.end()); }

// :14-22
Tile op32_potrf(const Tile &A, int N)
{
    auto a = to32(*A);
    blas().spotrf2(kRowMajor, 'L', N, a.data(), N);
    return make_tile(from32(a));
}

// :24-52
Tile op32_trsm(const Tile &L, const Tile &A, int N, int M, int trans, int side)
{
    auto l = to32(*L), a = to32(*A);
    blas().strsm(kRowMajor, side, kLower, trans, kNonUnit, N, M, 1.0f, l.data(), N, a.data(), M);
    return make_tile(from32(a));
}

// :54-65: A = A - B B^T (lower), all in float
Tile op32_syrk(const Tile &A, const Tile &B, int N)
{
    auto a = to32(*A), b = to32(*B);
    blas().ssyrk(kRowMajor, kLower, kNoTrans, N, N, -1.0f, b.data(), N, 1.0f, a.data(), N);
    return make_tile(from32(a));
}

// :67-100: C = C - op(A) op(B), all in float
Tile op32_gemm(const Tile &A, const Tile &B, const Tile &C, int N, int M, int K, int tA, int tB)
{
    auto a = to32(*A), b = to32(*B), c = to32(*C);
    blas().sgemm(kRowMajor, tA, tB, K, M, N, -1.0f, a.data(), K, b.data(), M, 1.0f, c.data(), M);
    return make_tile(from32(c));
}

// The mixed-precision trailing update: the FP32 adapter's product of float
// copies (FP32 sums) subtracted from the FP64 tile.
Tile op_syrk_f32prod(const Tile &A, const Tile &B, int N)
{
    auto b = to32(*B);
    std::vector<float> c(static_cast<std::size_t>(N) * N, 0.0f);
    blas().ssyrk(kRowMajor, kLower, kNoTrans, N, N, 1.0f, b.data(), N, 0.0f, c.data(), N);
    Vec a = *A;
    for (std::size_t i = 0; i < a.size(); i++)
        a[i] -= static_cast<double>(c[i]);
    return make_tile(std::move(a));
}

Tile op_gemm_f32prod(const Tile &A, const Tile &B, const Tile &C, int N, int M, int K, int tA, int tB)
{
    auto a = to32(*A), b = to32(*B);
    std::vector<float> p(static_cast<std::size_t>(K) * M, 0.0f);
    blas().sgemm(kRowMajor, tA, tB, K, M, N, 1.0f, a.data(), K, b.data(), M, 0.0f, p.data(), M);
    Vec c = *C;
    for (std::size_t i = 0; i < c.size(); i++)
        c[i] -= static_cast<double>(p[i]);
    return make_tile(std::move(c));
}

// core/src/cpu/gp_uncertainty.cpp:6-19
Tile op_diagonal(const Tile &A, std::size_t M)
{
    Vec t;
    t.reserve(M);
    for (std::size_t i = 0; i < M; ++i)
        t.push_back((*A)[i * M + i]);
    return make_tile(std::move(t));
}

// gp_optimizer.cpp:187-203: y^T alpha + sum_i log(L_ii^2) over one diagonal tile
double tile_loss(const Vec &Kdiag, const Vec &alpha, const Vec &y, std::size_t N)
{
    double l = op_dot(y, alpha, static_cast<int>(N));
    for (std::size_t i = 0; i < N; i++)
    {
        double d = Kdiag[i * N + i];
        l += std::log(d * d);
    }
    return l;
}

// gp_optimizer.cpp:205-218
double sum_losses(const std::vector<double> &losses, std::size_t N, std::size_t n_tiles)
{
    double l = 0.0;
    double Nn = static_cast<double>(N * n_tiles);
    for (std::size_t i = 0; i < n_tiles; i++)
        l += losses[i];
    l += Nn * std::log(2.0 * std::numbers::pi);
    return 0.5 * l / Nn;
}

// ---------------------------------------------------------------------------
// Tiled algorithms: core/src/cpu/tiled_algorithms.cpp:14-448

// :14-57
void cholesky_tiled(Tiles &A, int N, std::size_t T)
{
    for (std::size_t k = 0; k < T; k++)
    {
        A[k * T + k] = after([N](const Tile &a) { return op_potrf(a, N); }, A[k * T + k]);
        for (std::size_t m = k + 1; m < T; m++)
            A[m * T + k] = after([N](const Tile &l, const Tile &a) { return op_trsm(l, a, N, N, kTrans, kRight); },
                                 A[k * T + k],
                                 A[m * T + k]);
        for (std::size_t m = k + 1; m < T; m++)
        {
            A[m * T + m] =
                after([N](const Tile &a, const Tile &b) { return op_syrk(a, b, N); }, A[m * T + m], A[m * T + k]);
            for (std::size_t n = k + 1; n < m; n++)
                A[m * T + n] = after([N](const Tile &a, const Tile &b, const Tile &c)
                                     { return op_gemm(a, b, c, N, N, N, kNoTrans, kTrans); },
                                     A[m * T + k],
                                     A[n * T + k],
                                     A[m * T + n]);
        }
    }
}

// The DAG of cholesky_tiled with reduced-precision tile ops: level 1 (the
// mixed-precision comparator) computes the SYRK/GEMM trailing products in
// FP32 and subtracts them from the FP64 tiles; level 2 runs all four ops
// through the FP32 adapter.
void cholesky_tiled_reduced(Tiles &A, int N, std::size_t T, int level)
{
    const bool all32 = level >= 2;
    for (std::size_t k = 0; k < T; k++)
    {
        A[k * T + k] = after([N, all32](const Tile &a) { return all32 ? op32_potrf(a, N) : op_potrf(a, N); }, A[k * T + k]);
        for (std::size_t m = k + 1; m < T; m++)
            A[m * T + k] = after([N, all32](const Tile &l, const Tile &a)
                                 { return all32 ? op32_trsm(l, a, N, N, kTrans, kRight) : op_trsm(l, a, N, N, kTrans, kRight); },
                                 A[k * T + k],
                                 A[m * T + k]);
        for (std::size_t m = k + 1; m < T; m++)
        {
            A[m * T + m] = after([N, all32](const Tile &a, const Tile &b) { return all32 ? op32_syrk(a, b, N) : op_syrk_f32prod(a, b, N); },
                                 A[m * T + m],
                                 A[m * T + k]);
            for (std::size_t n = k + 1; n < m; n++)
                A[m * T + n] = after(
                    [N, all32](const Tile &a, const Tile &b, const Tile &c)
                    { return all32 ? op32_gemm(a, b, c, N, N, N, kNoTrans, kTrans) : op_gemm_f32prod(a, b, c, N, N, N, kNoTrans, kTrans); },
                    A[m * T + k],
                    A[n * T + k],
                    A[m * T + n]);
        }
    }
}

// :61-86
void forward_solve(Tiles &L, Tiles &b, int N, std::size_t T)
{
    for (std::size_t k = 0; k < T; k++)
    {
        b[k] = after([N](const Tile &l, const Tile &x) { return op_trsv(l, x, N, kNoTrans); }, L[k * T + k], b[k]);
        for (std::size_t m = k + 1; m < T; m++)
            b[m] = after([N](const Tile &a, const Tile &x, const Tile &y) { return op_gemv(a, x, y, N, N, -1.0, kNoTrans); },
                         L[m * T + k],
                         b[k],
                         b[m]);
    }
}

// :88-115
void backward_solve(Tiles &L, Tiles &b, int N, std::size_t T)
{
    for (int k_ = static_cast<int>(T) - 1; k_ >= 0; k_--)
    {
        std::size_t k = static_cast<std::size_t>(k_);
        b[k] = after([N](const Tile &l, const Tile &x) { return op_trsv(l, x, N, kTrans); }, L[k * T + k], b[k]);
        for (int m_ = k_ - 1; m_ >= 0; m_--)
        {
            std::size_t m = static_cast<std::size_t>(m_);
            b[m] = after([N](const Tile &a, const Tile &x, const Tile &y) { return op_gemv(a, x, y, N, N, -1.0, kTrans); },
                         L[k * T + m],
                         b[k],
                         b[m]);
        }
    }
}

// :117-149
void forward_solve_matrix(Tiles &L, Tiles &B, int N, int M, std::size_t T, std::size_t MT)
{
    for (std::size_t c = 0; c < MT; c++)
        for (std::size_t k = 0; k < T; k++)
        {
            B[k * MT + c] = after([N, M](const Tile &l, const Tile &a) { return op_trsm(l, a, N, M, kNoTrans, kLeft); },
                                  L[k * T + k],
                                  B[k * MT + c]);
            for (std::size_t m = k + 1; m < T; m++)
                B[m * MT + c] = after([N, M](const Tile &a, const Tile &b, const Tile &cc)
                                      { return op_gemm(a, b, cc, N, M, N, kNoTrans, kNoTrans); },
                                      L[m * T + k],
                                      B[k * MT + c],
                                      B[m * MT + c]);
        }
}

// :151-185
void backward_solve_matrix(Tiles &L, Tiles &B, int N, int M, std::size_t T, std::size_t MT)
{
    for (std::size_t c = 0; c < MT; c++)
        for (int k_ = static_cast<int>(T) - 1; k_ >= 0; k_--)
        {
            std::size_t k = static_cast<std::size_t>(k_);
            B[k * MT + c] = after([N, M](const Tile &l, const Tile &a) { return op_trsm(l, a, N, M, kTrans, kLeft); },
                                  L[k * T + k],
                                  B[k * MT + c]);
            for (int m_ = k_ - 1; m_ >= 0; m_--)
            {
                std::size_t m = static_cast<std::size_t>(m_);
                B[m * MT + c] = after([N, M](const Tile &a, const Tile &b, const Tile &cc)
                                      { return op_gemm(a, b, cc, N, M, N, kTrans, kNoTrans); },
                                      L[k * T + m],
                                      B[k * MT + c],
                                      B[m * MT + c]);
            }
        }
}

// :187-210 rhs[k] += A[k][m] x[m], m ascending
void matrix_vector(Tiles &A, Tiles &x, Tiles &rhs, int Nr, int Nc, std::size_t T, std::size_t MT)
{
    for (std::size_t k = 0; k < MT; k++)
        for (std::size_t m = 0; m < T; m++)
            rhs[k] = after([Nr, Nc](const Tile &a, const Tile &v, const Tile &y)
                           { return op_gemv(a, v, y, Nr, Nc, 1.0, kNoTrans); },
                           A[k * T + m],
                           x[m],
                           rhs[k]);
}

// :212-228
void sym_matrix_matrix_diag(Tiles &V, Tiles &w, int N, int M, std::size_t T, std::size_t MT)
{
    for (std::size_t i = 0; i < MT; ++i)
        for (std::size_t n = 0; n < T; ++n)
            w[i] = after([N, M](const Tile &a, const Tile &r) { return op_dot_diag_syrk(a, r, N, M); }, V[n * MT + i], w[i]);
}

// :230-254
void sym_matrix_matrix(Tiles &V, Tiles &S, int N, int M, std::size_t T, std::size_t MT)
{
    for (std::size_t c = 0; c < MT; c++)
        for (std::size_t k = 0; k < MT; k++)
            for (std::size_t m = 0; m < T; m++)
                S[c * MT + k] = after([N, M](const Tile &a, const Tile &b, const Tile &cc)
                                      { return op_gemm(a, b, cc, N, M, M, kTrans, kNoTrans); },
                                      V[m * MT + c],
                                      V[m * MT + k],
                                      S[c * MT + k]);
}

// :256-263
void vector_difference(Tiles &minuend, Tiles &sub, int M, std::size_t MT)
{
    for (std::size_t i = 0; i < MT; i++)
        sub[i] = after([M](const Tile &y, const Tile &x) { return op_axpy(y, x, M); }, minuend[i], sub[i]);
}

// :265-272
void matrix_diagonal(Tiles &A, Tiles &v, int M, std::size_t MT)
{
    for (std::size_t i = 0; i < MT; i++)
        v[i] = after([M](const Tile &a) { return op_diagonal(a, static_cast<std::size_t>(M)); }, A[i * MT + i]);
}

// :274-294
FScalar loss_tiled(Tiles &L, Tiles &alpha, Tiles &y, int N, std::size_t T)
{
    std::vector<FScalar> parts;
    for (std::size_t k = 0; k < T; k++)
        parts.push_back(after([N](const Tile &l, const Tile &a, const Tile &yy)
                              { return tile_loss(*l, *a, *yy, static_cast<std::size_t>(N)); },
                              L[k * T + k],
                              alpha[k],
                              y[k]));
    return after_all([N, T](const std::vector<double> &v) { return sum_losses(v, static_cast<std::size_t>(N), T); },
                     parts);
}

// :296-448. Gradient of the loss w.r.t. one hyperparameter (in softplus
// space), then the Adam moment updates and step on `sek`.
void update_hyperparameter(const Tiles &Kinv,
                           const Tiles &gradK,
                           const Tiles &alpha,
                           const AdamParams &adam,
                           SEKParams &sek,
                           int N,
                           std::size_t T,
                           std::size_t iter,
                           std::size_t idx)
{
    FScalar trace = FScalar::now(0.0);
    FScalar dotv = FScalar::now(0.0);
    bool jitter = false;
    double factor = 1.0;
    if (idx == 0 || idx == 1)
    {
        Tiles diag, inter;
        for (std::size_t d = 0; d < T; d++)
        {
            diag.push_back(spawn([N] { return tile_zeros(static_cast<std::size_t>(N)); }));
            inter.push_back(spawn([N] { return tile_zeros(static_cast<std::size_t>(N)); }));
        }
        // diag(K^-1 dK) accumulated over j ascending (tiled_algorithms.cpp:344-356)
        for (std::size_t i = 0; i < T; ++i)
            for (std::size_t j = 0; j < T; ++j)
                diag[i] = after([N](const Tile &a, const Tile &b, const Tile &r) { return op_dot_diag_gemm(a, b, r, N, N); },
                                Kinv[i * T + j],
                                gradK[j * T + i],
                                diag[i]);
        for (std::size_t j = 0; j < T; ++j)
            trace = after([](const Tile &d, double t) { return t + std::reduce(d->begin(), d->end()); }, diag[j], trace);
        // alpha^T dK alpha (tiled_algorithms.cpp:366-388)
        for (std::size_t k = 0; k < T; k++)
            for (std::size_t m = 0; m < T; m++)
                inter[k] = after([N](const Tile &a, const Tile &x, const Tile &y) { return op_gemv(a, x, y, N, N, 1.0, kNoTrans); },
                                 gradK[k * T + m],
                                 alpha[m],
                                 inter[k]);
        for (std::size_t j = 0; j < T; ++j)
            dotv = after([](const Tile &v, const Tile &a, double r)
                         { return r + op_dot(*v, *a, static_cast<int>(a->size())); },
                         inter[j],
                         alpha[j],
                         dotv);
    }
    else if (idx == 2)
    {
        jitter = true;
        for (std::size_t j = 0; j < T; ++j)
            trace = after(
                [N](const Tile &t, double tr)
                {
                    double local = 0.0;
                    for (std::size_t i = 0; i < static_cast<std::size_t>(N); ++i)
                        local += (*t)[i * static_cast<std::size_t>(N) + i];
                    return tr + local;
                },
                Kinv[j * T + j],
                trace);
        for (std::size_t j = 0; j < T; ++j)
            dotv = after([](const Tile &a, const Tile &b, double r) { return r + op_dot(*a, *b, static_cast<int>(b->size())); },
                         alpha[j],
                         alpha[j],
                         dotv);
        factor = compute_sigmoid(to_unconstrained(sek.noise_variance, true));
    }
    else
    {
        throw std::invalid_argument("Invalid param_idx");
    }
    // gp_optimizer.cpp:222-225
    double gradient = factor * (0.5 / static_cast<double>(static_cast<std::size_t>(N) * T) * (trace.get() - dotv.get()));
    sek.m_T[idx] = first_moment(gradient, sek.m_T[idx], adam.beta1);
    sek.w_T[idx] = second_moment(gradient, sek.w_T[idx], adam.beta2);
    double u = to_unconstrained(sek.get_param(idx), jitter);
    double updated = adam_update(u, adam, sek.m_T[idx], sek.w_T[idx], iter);
    sek.set_param(idx, to_constrained(updated, jitter));
}

// Lower K tiles of the covariance (gp_functions.cpp:40-53 and repeats).
Tiles assemble_K(const Vec &x, const SEKParams &sek, std::size_t T, int N, int r)
{
    Tiles K(T * T);
    for (std::size_t i = 0; i < T; i++)
        for (std::size_t j = 0; j <= i; j++)
            K[i * T + j] = spawn([=, &x] { return tile_covariance(i, j, N, r, sek, x); });
    return K;
}

Tiles assemble_y(const Vec &y, std::size_t T, int N)
{
    if (y.size() < T * static_cast<std::size_t>(N))
        throw std::invalid_argument("oracle: training output shorter than n_tiles * n_tile_size");
    Tiles t;
    for (std::size_t i = 0; i < T; i++)
        t.push_back(spawn([=, &y] { return tile_output(i, N, y); }));
    return t;
}

Tiles assemble_cross(const Vec &z, const Vec &x, const SEKParams &sek, std::size_t MT, std::size_t T, int M, int N, int r)
{
    Tiles c;
    for (std::size_t i = 0; i < MT; i++)
        for (std::size_t j = 0; j < T; j++)
            c.push_back(spawn([=, &z, &x] { return tile_cross(i, j, M, N, r, sek, z, x); }));
    return c;
}

Tiles zeros(std::size_t count, int N)
{
    Tiles t;
    for (std::size_t i = 0; i < count; i++)
        t.push_back(spawn([N] { return tile_zeros(static_cast<std::size_t>(N)); }));
    return t;
}

std::vector<double> flatten(const Tiles &t)
{
    std::vector<double> out;
    for (auto &f : t)
    {
        const Tile &v = f.get();
        out.insert(out.end(), v->begin(), v->end());
    }
    return out;
}

void check_input(const Vec &x, std::size_t n, int r, const char *what)
{
    if (r < 1)
        throw std::invalid_argument("oracle: n_regressors must be >= 1");
    if (n > 0 && x.size() < n + static_cast<std::size_t>(r) - 1)
        throw std::invalid_argument(std::string("oracle: ") + what + " shorter than n + n_reg - 1");
}

// One optimizer iteration's graph (gp_functions.cpp:867-1026 / :1109-1268).
double optimize_iteration(const Vec &x,
                          const Tiles &y_tiles,
                          int T_,
                          int N,
                          int r,
                          const AdamParams &adam,
                          SEKParams &sek,
                          const std::vector<bool> &trainable,
                          std::size_t iter)
{
    const std::size_t T = static_cast<std::size_t>(T_);
    Tiles K(T * T), gl(T * T), gv(T * T), alpha(T), Kinv(T * T);
    const SEKParams snap = sek;  // dataflow copies the parameters at launch
    for (std::size_t i = 0; i < T; i++)
        for (std::size_t j = 0; j <= i; j++)
        {
            FTile D = spawn([=, &x] { return tile_distance(i, j, N, r, snap, x); });
            K[i * T + j] = after([=](const Tile &d) { return make_tile(tile_covariance_from_distance(i, j, N, snap, *d)); }, D);
            if (trainable[0])
            {
                gl[i * T + j] = after([=](const Tile &d) { return make_tile(tile_grad_l(N, snap, *d)); }, D);
                if (i != j)
                    gl[j * T + i] = after([=](const Tile &g) { return make_tile(tile_transpose(N, N, *g)); }, gl[i * T + j]);
            }
            if (trainable[1])
            {
                gv[i * T + j] = after([=](const Tile &d) { return make_tile(tile_grad_v(N, snap, *d)); }, D);
                if (i != j)
                    gv[j * T + i] = after([=](const Tile &g) { return make_tile(tile_transpose(N, N, *g)); }, gv[i * T + j]);
            }
        }
    for (std::size_t i = 0; i < T; i++)
        alpha[i] = spawn([N] { return tile_zeros(static_cast<std::size_t>(N)); });
    for (std::size_t i = 0; i < T; i++)
        for (std::size_t j = 0; j < T; j++)
            Kinv[i * T + j] = (i == j) ? spawn([N] { return tile_identity(static_cast<std::size_t>(N)); })
                                       : spawn([N] { return tile_zeros(static_cast<std::size_t>(N) * N); });
    cholesky_tiled(K, N, T);
    forward_solve_matrix(K, Kinv, N, N, T, T);
    backward_solve_matrix(K, Kinv, N, N, T, T);
    Tiles ycopy = y_tiles;
    matrix_vector(Kinv, ycopy, alpha, N, N, T, T);
    FScalar loss = loss_tiled(K, alpha, ycopy, N, T);
    if (trainable[0])
        update_hyperparameter(Kinv, gl, alpha, adam, sek, N, T, iter, 0);
    if (trainable[1])
        update_hyperparameter(Kinv, gv, alpha, adam, sek, N, T, iter, 1);
    if (trainable[2])
        update_hyperparameter(Kinv, Tiles{}, alpha, adam, sek, N, T, iter, 2);
    return loss.get();
}

}  // namespace

// ---------------------------------------------------------------------------
void set_threads(unsigned n)
{
    std::lock_guard<std::mutex> g(g_pool_mu);
    g_pool.reset();
    g_pool = std::make_unique<Pool>(n ? n : 1);
}

unsigned threads() { return pool().size(); }

std::string blas_name() { return blas().name; }

// ---------------------------------------------------------------------------
// API calls: core/src/cpu/gp_functions.cpp

Matrix cholesky(const Vec &x, const SEKParams &sek, int T_, int N, int r)  // :17-78
{
    const std::size_t T = static_cast<std::size_t>(T_);
    check_input(x, T * static_cast<std::size_t>(N), r, "training input");
    Tiles K = assemble_K(x, sek, T, N, r);
    cholesky_tiled(K, N, T);
    Matrix result(T * T);
    for (std::size_t i = 0; i < T; i++)
        for (std::size_t j = 0; j <= i; j++)
            result[i * T + j] = *K[i * T + j].get();
    return result;
}

std::vector<double> predict(const Vec &x, const Vec &y, const Vec &z, const SEKParams &sek, int T_, int N, int MT_, int M, int r)
{  // :80-216
    const std::size_t T = static_cast<std::size_t>(T_), MT = static_cast<std::size_t>(MT_);
    check_input(x, T * static_cast<std::size_t>(N), r, "training input");
    check_input(z, MT * static_cast<std::size_t>(M), r, "test input");
    Tiles K = assemble_K(x, sek, T, N, r);
    Tiles alpha = assemble_y(y, T, N);
    Tiles cross = assemble_cross(z, x, sek, MT, T, M, N, r);
    Tiles pred = zeros(MT, M);
    cholesky_tiled(K, N, T);
    forward_solve(K, alpha, N, T);
    backward_solve(K, alpha, N, T);
    matrix_vector(cross, alpha, pred, M, N, T, MT);
    return flatten(pred);
}

Matrix predict_with_uncertainty(const Vec &x, const Vec &y, const Vec &z, const SEKParams &sek, int T_, int N, int MT_, int M, int r)
{  // :218-450
    const std::size_t T = static_cast<std::size_t>(T_), MT = static_cast<std::size_t>(MT_);
    check_input(x, T * static_cast<std::size_t>(N), r, "training input");
    check_input(z, MT * static_cast<std::size_t>(M), r, "test input");
    Tiles K = assemble_K(x, sek, T, N, r);
    Tiles alpha = assemble_y(y, T, N);
    Tiles cross = assemble_cross(z, x, sek, MT, T, M, N, r);
    Tiles pred = zeros(MT, M);
    Tiles prior;
    for (std::size_t i = 0; i < MT; i++)
        prior.push_back(spawn([=, &z] { return tile_prior_diag(i, i, M, r, sek, z); }));
    Tiles crossT;  // index j * MT + i
    for (std::size_t j = 0; j < T; j++)
        for (std::size_t i = 0; i < MT; i++)
            crossT.push_back(after([M, N](const Tile &c) { return make_tile(tile_transpose(M, N, *c)); }, cross[i * T + j]));
    Tiles unc = zeros(MT, M);
    cholesky_tiled(K, N, T);
    forward_solve(K, alpha, N, T);
    backward_solve(K, alpha, N, T);
    matrix_vector(cross, alpha, pred, M, N, T, MT);
    forward_solve_matrix(K, crossT, N, M, T, MT);
    sym_matrix_matrix_diag(crossT, unc, N, M, T, MT);
    vector_difference(prior, unc, M, MT);
    return { flatten(pred), flatten(unc) };
}

Matrix predict_with_full_cov(const Vec &x, const Vec &y, const Vec &z, const SEKParams &sek, int T_, int N, int MT_, int M, int r)
{  // :452-696
    const std::size_t T = static_cast<std::size_t>(T_), MT = static_cast<std::size_t>(MT_);
    check_input(x, T * static_cast<std::size_t>(N), r, "training input");
    check_input(z, MT * static_cast<std::size_t>(M), r, "test input");
    Tiles K = assemble_K(x, sek, T, N, r);
    Tiles alpha = assemble_y(y, T, N);
    Tiles cross = assemble_cross(z, x, sek, MT, T, M, N, r);
    Tiles pred = zeros(MT, M);
    Tiles prior(MT * MT);
    for (std::size_t i = 0; i < MT; i++)
        for (std::size_t j = 0; j <= i; j++)
        {
            prior[i * MT + j] = spawn([=, &z] { return tile_full_prior(i, j, M, r, sek, z); });
            if (i != j)
                prior[j * MT + i] = after([M](const Tile &p) { return make_tile(tile_transpose(M, M, *p)); }, prior[i * MT + j]);
        }
    Tiles crossT;
    for (std::size_t j = 0; j < T; j++)
        for (std::size_t i = 0; i < MT; i++)
            crossT.push_back(after([M, N](const Tile &c) { return make_tile(tile_transpose(M, N, *c)); }, cross[i * T + j]));
    Tiles unc = zeros(MT, M);
    cholesky_tiled(K, N, T);
    forward_solve(K, alpha, N, T);
    backward_solve(K, alpha, N, T);
    forward_solve_matrix(K, crossT, N, M, T, MT);
    matrix_vector(cross, alpha, pred, M, N, T, MT);
    sym_matrix_matrix(crossT, prior, N, M, T, MT);
    matrix_diagonal(prior, unc, M, MT);
    return { flatten(pred), flatten(unc) };
}

double compute_loss(const Vec &x, const Vec &y, const SEKParams &sek, int T_, int N, int r)  // :700-781
{
    const std::size_t T = static_cast<std::size_t>(T_);
    check_input(x, T * static_cast<std::size_t>(N), r, "training input");
    Tiles K = assemble_K(x, sek, T, N, r);
    Tiles yt = assemble_y(y, T, N);
    Tiles alpha = assemble_y(y, T, N);
    cholesky_tiled(K, N, T);
    forward_solve(K, alpha, N, T);
    backward_solve(K, alpha, N, T);
    return loss_tiled(K, alpha, yt, N, T).get();
}

// alpha = K~^-1 y from a reduced-precision factor, refined in FP64: r = y -
// K alpha with the FP64 covariance tiles (lower tiles and their transposes),
// delta from the factor, alpha += delta while the corrections shrink, until
// a correction is at most 1e-12 max|alpha|. Returns the steps taken.
int refine_alpha_reduced(Tiles &L, Tiles &K, Tiles &alpha, const Tiles &y, int N, std::size_t T)
{
    double prev = 1e300;
    int steps = 0;
    for (int it = 0; it < 30; it++)
    {
        Tiles r;
        for (std::size_t i = 0; i < T; i++)
            r.push_back(y[i]);
        for (std::size_t i = 0; i < T; i++)
            for (std::size_t j = 0; j < T; j++)
            {
                if (j <= i)
                    r[i] = after([N](const Tile &a, const Tile &x, const Tile &b) { return op_gemv(a, x, b, N, N, -1.0, kNoTrans); },
                                 K[i * T + j], alpha[j], r[i]);
                else
                    r[i] = after([N](const Tile &a, const Tile &x, const Tile &b) { return op_gemv(a, x, b, N, N, -1.0, kTrans); },
                                 K[j * T + i], alpha[j], r[i]);
            }
        forward_solve(L, r, N, T);
        backward_solve(L, r, N, T);
        double dmax = 0.0, amax = 0.0;
        for (std::size_t i = 0; i < T; i++)
        {
            const Vec &d = *r[i].get();
            const Vec &a = *alpha[i].get();
            for (std::size_t e = 0; e < d.size(); e++)
            {
                dmax = std::max(dmax, std::fabs(d[e]));
                amax = std::max(amax, std::fabs(a[e]));
            }
        }
        if (it > 0 && dmax >= prev)
            break;
        for (std::size_t i = 0; i < T; i++)
        {
            Vec a = *alpha[i].get();
            const Vec &d = *r[i].get();
            for (std::size_t e = 0; e < a.size(); e++)
                a[e] += d[e];
            alpha[i] = Flow<Tile>::now(make_tile(std::move(a)));
        }
        steps = it + 1;
        if (dmax <= 1e-12 * amax || !(dmax < 0.5 * prev))
            break;
        prev = dmax;
    }
    return steps;
}

Reduced predict_with_uncertainty_reduced(const Vec &x, const Vec &y, const Vec &z, const SEKParams &sek, int T_, int N, int MT_, int M,
                                         int r, int level)
{
    const std::size_t T = static_cast<std::size_t>(T_), MT = static_cast<std::size_t>(MT_);
    check_input(x, T * static_cast<std::size_t>(N), r, "training input");
    check_input(z, MT * static_cast<std::size_t>(M), r, "test input");
    Tiles K = assemble_K(x, sek, T, N, r);
    Tiles K64 = K;  // tiles are immutable: the factorisation replaces K's entries
    Tiles yt = assemble_y(y, T, N);
    Tiles alpha = assemble_y(y, T, N);
    Tiles cross = assemble_cross(z, x, sek, MT, T, M, N, r);
    Tiles pred = zeros(MT, M);
    Tiles prior;
    for (std::size_t i = 0; i < MT; i++)
        prior.push_back(spawn([=, &z] { return tile_prior_diag(i, i, M, r, sek, z); }));
    Tiles crossT;
    for (std::size_t j = 0; j < T; j++)
        for (std::size_t i = 0; i < MT; i++)
            crossT.push_back(after([M, N](const Tile &c) { return make_tile(tile_transpose(M, N, *c)); }, cross[i * T + j]));
    Tiles unc = zeros(MT, M);
    if (level <= 0)
        cholesky_tiled(K, N, T);
    else
        cholesky_tiled_reduced(K, N, T, level);
    forward_solve(K, alpha, N, T);
    backward_solve(K, alpha, N, T);
    Reduced out;
    out.refine_steps = refine_alpha_reduced(K, K64, alpha, yt, N, T);
    matrix_vector(cross, alpha, pred, M, N, T, MT);
    forward_solve_matrix(K, crossT, N, M, T, MT);
    sym_matrix_matrix_diag(crossT, unc, N, M, T, MT);
    vector_difference(prior, unc, M, MT);
    out.mean = flatten(pred);
    out.var = flatten(unc);
    out.loss = loss_tiled(K, alpha, yt, N, T).get();
    return out;
}

std::vector<double> optimize(const Vec &x,
                             const Vec &y,
                             int T,
                             int N,
                             int r,
                             const AdamParams &adam,
                             SEKParams &sek,
                             const std::vector<bool> &trainable)
{  // :783-1030
    check_input(x, static_cast<std::size_t>(T) * static_cast<std::size_t>(N), r, "training input");
    if (trainable.size() < 3)
        throw std::invalid_argument("oracle: trainable needs 3 entries");
    Tiles yt = assemble_y(y, static_cast<std::size_t>(T), N);
    std::vector<double> losses;
    for (std::size_t iter = 0; iter < static_cast<std::size_t>(adam.opt_iter); iter++)
        losses.push_back(optimize_iteration(x, yt, T, N, r, adam, sek, trainable, iter));
    return losses;
}

double optimize_step(const Vec &x,
                     const Vec &y,
                     int T,
                     int N,
                     int r,
                     AdamParams &adam,
                     SEKParams &sek,
                     const std::vector<bool> &trainable,
                     int iter)
{  // :1032-1269
    check_input(x, static_cast<std::size_t>(T) * static_cast<std::size_t>(N), r, "training input");
    if (trainable.size() < 3)
        throw std::invalid_argument("oracle: trainable needs 3 entries");
    Tiles yt = assemble_y(y, static_cast<std::size_t>(T), N);
    return optimize_iteration(x, yt, T, N, r, adam, sek, trainable, static_cast<std::size_t>(iter));
}

}  // namespace oracle
