// Host orchestration of the gfx950 GP backend.
//
// Data layout in HBM: the training covariance is one dense row-major
// n_pad x lda matrix (n_pad = n rounded up to 128) whose lower triangle is
// overwritten by the Cholesky factor; rows/columns >= n hold an identity
// block. Vectors are zero padded to n_pad. The tile split of the reference
// API (n_tiles x n_tile_size) only shapes inputs and the cholesky() output;
// the factorisation uses its own blocking.
#include "engine.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numbers>

namespace gpamd
{

[[noreturn]] void throw_hip(hipError_t e, const char *expr, const char *file, int line)
{
    const int status = (e == hipErrorOutOfMemory) ? GPRAT_ERR_OUT_OF_MEMORY : GPRAT_ERR_HIP;
    throw Error(status,
                std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ") in " + expr + " at "
                    + file + ":" + std::to_string(line));
}

// ---------------------------------------------------------------------------
// DevicePool

DevicePool::~DevicePool()
{
    hipSetDevice(device_);
    for (auto &kv : free_)
    {
        if (kv.second.returned)
            hipEventDestroy(kv.second.returned);
        hipFree(kv.second.p);
    }
    for (auto e : spare_)
        hipEventDestroy(e);
    for (auto &kv : live_)
        hipFree(kv.first);
}

static std::size_t size_class(std::size_t bytes)
{
    const std::size_t g = bytes < (std::size_t(1) << 20) ? 4096 : (std::size_t(2) << 20);
    return ((bytes + g - 1) / g) * g;
}

void DevicePool::share_with(const std::vector<DevicePool *> &siblings, hipStream_t stream)
{
    std::lock_guard<std::mutex> g(mu_);
    siblings_.clear();
    for (DevicePool *s : siblings)
        if (s != this)
            siblings_.push_back(s);
    stream_ = stream;
}

bool DevicePool::take(std::size_t want, Cached &out, std::size_t &size)
{
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.lower_bound(want);
    // reuse a cached block if it is not more than 25% larger than needed
    if (it == free_.end() || it->first > want + want / 4)
        return false;
    out = it->second;
    size = it->first;
    cached_ -= it->first;
    free_.erase(it);
    return true;
}

void *DevicePool::get(std::size_t bytes)
{
    const std::size_t want = size_class(bytes ? bytes : 1);
    Cached b{};
    std::size_t size = 0;
    // own cache: this pool's stream order already covers the block's earlier uses
    bool found = take(want, b, size);
    bool foreign = false;
    for (std::size_t i = 0; !found && i < siblings_.size(); i++)
        foreign = found = siblings_[i]->take(want, b, size);
    if (found)
    {
        if (b.returned && foreign)
            GP_HIP(hipEventSynchronize(b.returned));
        std::lock_guard<std::mutex> g(mu_);
        if (b.returned)
            spare_.push_back(b.returned);
        live_[b.p] = size;
        borrowed_ += foreign ? 1 : 0;
        return b.p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess)
    {
        (void)hipGetLastError();
        if (on_out_of_memory)
            on_out_of_memory();
        trim(0);
        if (hipMalloc(&p, want) != hipSuccess)
        {
            (void)hipGetLastError();
            throw Error(GPRAT_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(want) + " bytes failed");
        }
    }
    std::lock_guard<std::mutex> g(mu_);
    live_[p] = want;
    return p;
}

void DevicePool::put(void *p)
{
    if (!p)
        return;
    hipEvent_t ev = nullptr;
    hipStream_t stream = nullptr;
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = live_.find(p);
        if (it == live_.end())
            return;
        stream = stream_;
        if (stream && !spare_.empty())
        {
            ev = spare_.back();
            spare_.pop_back();
        }
    }
    if (stream)
    {
        // the event must belong to the stream's device, whichever device
        // the releasing thread has current
        int cur = -1;
        GP_HIP(hipGetDevice(&cur));
        if (cur != device_)
            GP_HIP(hipSetDevice(device_));
        if (!ev)
            GP_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        GP_HIP(hipEventRecord(ev, stream));
        if (cur != device_)
            GP_HIP(hipSetDevice(cur));
    }
    std::lock_guard<std::mutex> g(mu_);
    auto it = live_.find(p);
    free_.emplace(it->second, Cached{ p, ev });
    cached_ += it->second;
    live_.erase(it);
}

void DevicePool::release_cached() { trim(0); }

void DevicePool::trim(std::size_t keep)
{
    // the blocks leave the cache under the lock (largest first); the wait
    // for kernels that may still use them and the frees happen after it is
    // released, so put() and get() from other threads never wait on the device
    std::vector<Cached> drop;
    {
        std::lock_guard<std::mutex> g(mu_);
        while (cached_ > keep && !free_.empty())
        {
            auto it = std::prev(free_.end());
            drop.push_back(it->second);
            cached_ -= it->first;
            free_.erase(it);
        }
    }
    if (drop.empty())
        return;
    hipSetDevice(device_);
    hipDeviceSynchronize();
    for (const Cached &b : drop)
    {
        if (b.returned)
            hipEventDestroy(b.returned);
        hipFree(b.p);
    }
}

template <class T>
DevBuf<T>::DevBuf(Context &c, std::size_t count) : pool_(&c.pool), n_(count)
{
    p_ = static_cast<T *>(pool_->get(count * sizeof(T)));
}

template <class T>
void DevBuf<T>::reset()
{
    if (p_ && pool_)
        pool_->put(p_);
    p_ = nullptr;
    n_ = 0;
}

template class DevBuf<double>;
template class DevBuf<float>;
template class DevBuf<int>;
template class DevBuf<int2>;

// ---------------------------------------------------------------------------
// Context

Context::Context(int dev, int ns) : device(dev), n_streams(ns), pool(dev)
{
    make_current();
    GP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    GP_HIP(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
    panel = trail = main;
    // one stream: the panel factorisation runs in line (no look-ahead)
    if (ns >= 2)
        GP_HIP(hipStreamCreateWithFlags(&panel, hipStreamNonBlocking));
    info = DevBuf<int>(*this, 1);
    scalars = DevBuf<double>(*this, 64);
    k::noalloc_scratch_init(device);
}

Context::~Context()
{
    hipSetDevice(device);
    hipDeviceSynchronize();
    for (auto &p : pending)
    {
        hipEventDestroy(p.begin);
        hipEventDestroy(p.end);
    }
    for (auto e : spare_events)
        hipEventDestroy(e);
    panel_plans.clear();
    batch_graphs.clear();
    info.reset();
    scalars.reset();
    if (trail && trail != main)
        hipStreamDestroy(trail);
    if (panel && panel != main)
        hipStreamDestroy(panel);
    if (main)
        hipStreamDestroy(main);
}

void Context::make_current() const { GP_HIP(hipSetDevice(device)); }

hipEvent_t Context::event()
{
    if (!spare_events.empty())
    {
        hipEvent_t e = spare_events.back();
        spare_events.pop_back();
        return e;
    }
    hipEvent_t e;
    GP_HIP(hipEventCreate(&e));
    return e;
}

void Context::recycle(hipEvent_t e) { spare_events.push_back(e); }

void Context::prof_begin(hipStream_t s, const char *family, Pending &p)
{
    p.family = family;
    p.begin = event();
    p.end = event();
    GP_HIP(hipEventRecord(p.begin, s));
}

void Context::prof_end(hipStream_t s, Pending &p)
{
    if (hipEventRecord(p.end, s) == hipSuccess)
        pending.push_back(p);
    else
    {
        (void)hipGetLastError();
        recycle(p.begin);
        recycle(p.end);
    }
}

void Context::resolve_profiling()
{
    for (auto &p : pending)
    {
        float ms = 0.0f;
        GP_HIP(hipEventElapsedTime(&ms, p.begin, p.end));
        KernelStat &st = stats[p.family];
        st.launches += 1;
        st.ms += ms;
        recycle(p.begin);
        recycle(p.end);
    }
    pending.clear();
}

void Context::sync()
{
    if (panel != main)
        GP_HIP(hipStreamSynchronize(panel));
    if (trail != main)
        GP_HIP(hipStreamSynchronize(trail));
    GP_HIP(hipStreamSynchronize(main));
    if (!pending.empty())
        resolve_profiling();
}

// Optional timing bracket around a launch sequence.


// ---------------------------------------------------------------------------
// Layout

long padded(long n) { return round_up(std::max<long>(n, 1), 128); }

long leading_dim(long n_pad)
{
    // Large power-of-two strides map consecutive rows to the same HBM
    // channels; offset them by half a KiB.
    if (n_pad >= 4096 && n_pad % 1024 == 0)
        return n_pad + 64;
    return n_pad;
}

// Look-ahead schedule (panel k+1 on its own stream while the trailing
// update of step k runs) or a single-stream schedule; GPRAT_LOOKAHEAD=0/1
// overrides the default.
static bool use_lookahead()
{
    if (const char *e = std::getenv("GPRAT_LOOKAHEAD"))
        return std::atoi(e) != 0;
    return true;
}

static int outer_block(int n_pad)
{
    if (const char *e = std::getenv("GPRAT_CHOLESKY_BLOCK"))
    {
        int v = std::atoi(e);
        if (v >= 128 && v % 128 == 0)
            return v;
    }
    // factorisation measured at n = 16384 / 32768: 512-wide panels
    // 35.6 / 210.6 ms, 768-wide 36.1 / 204.9 ms, 1024-wide 39.2 / 207.2 ms;
    // wider panels give the trailing update a longer K but lengthen the
    // diagonal-block chain and the panel stream's head work. The variance
    // solve steps through the same blocks, so a whole request at n = 16384
    // (m = 4096) favours 768 (56.5 against 57.1 ms with 512)
    return n_pad <= 4096 ? 256 : (n_pad < 12288 ? 512 : 768);
}

// Workgroup slots (out of two per CU) that look-ahead launches on the
// trailing stream leave free, so the panel stream's small kernels start at
// once instead of queueing behind a saturating update (no stream
// priorities); GPRAT_TRAILING_RESERVE overrides. Measured with 768-wide
// panels (request at n = 16384, alternating runs): 16 slots 56.5-56.8 ms,
// 24 slots 55.9-56.0 ms, 32 slots 55.9-56.0 ms; n = 32768 unchanged.
static int trailing_reserve()
{
    if (const char *e = std::getenv("GPRAT_TRAILING_RESERVE"))
        return std::max(0, std::atoi(e));
    return 24;
}

// Workgroups of the panel stream's head launches (the next diagonal
// block's rows of the panel and its update); GPRAT_HEAD_WORKGROUPS
// overrides. Request at n = 16384 (768-wide panels: 144 + 78 head tasks),
// alternating runs: 32 → 57.3-57.5 ms, 64 → 55.9-56.4, 96 → 55.3-55.7,
// 128 → 55.2-55.4, 160 → 55.7-55.8.
static int head_workgroups()
{
    if (const char *e = std::getenv("GPRAT_HEAD_WORKGROUPS"))
        return std::max(1, std::atoi(e));
    return 128;
}

// C[i][j] -= sum_k P[k][a_off + i] P[k][b_off + j] for i < M, j < N,
// i + diag_off >= j, with P the k-major copy of a factored panel (leading
// dimension ldp).
static void trailing_kmajor(hipStream_t s, const double *P, long ldp, long a_off, long b_off, long diag_off, double *C, long ldc, int M, int N,
                            int K, int max_wgs, int *queue)
{
    k::GemmArgs g{};
    g.A = P + a_off;
    g.lda = ldp;
    g.B = P + b_off;
    g.ldb = ldp;
    g.C = C;
    g.ldc = ldc;
    g.M = M;
    g.N = N;
    g.K = K;
    g.a_trans = true;
    g.b_trans = false;
    g.lower = true;
    g.diag_off = diag_off;
    g.role = k::GemmRole::TrailingT;
    g.max_wgs = max_wgs;
    g.queue = queue;
    k::gemm_f64(s, g);
}

// Diagonal-block task schedules (panel_plan.hpp), built once per shape and
// kept on the device for the context's lifetime.
const Context::PanelPlanDev &Context::panel_plan(int nb, int chain_wgs, int tail_wgs)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    chain_wgs = std::clamp(chain_wgs, 1, 4096);
    tail_wgs = std::clamp(tail_wgs, 1, 4096);
    const long key = (static_cast<long>(nb) << 26) | (static_cast<long>(chain_wgs) << 13) | tail_wgs;
    return cached_plan(key, [&] { return build_panel_plan(nb, chain_wgs, tail_wgs); });
}

const Context::PanelPlanDev &Context::head_plan(int nbh, int nb, int wgs)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    wgs = std::clamp(wgs, 1, 4096);
    const long key = (1L << 50) | (static_cast<long>(nbh) << 26) | (static_cast<long>(nb) << 13) | wgs;
    return cached_plan(key, [&] { return build_head_plan(nbh, nb, wgs); });
}

const Context::PanelPlanDev &Context::cached_plan(long key, const std::function<PanelPlan()> &build)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    auto it = panel_plans.find(key);
    if (it != panel_plans.end())
        return it->second;
    const PanelPlan p = build();
    PanelPlanDev d;
    d.tasks = DevBuf<int>(*this, p.tasks.size() * 4);
    d.wg_off = DevBuf<int>(*this, p.wg_off.size());
    make_current();
    GP_HIP(hipMemcpy(d.tasks.get(), p.tasks.data(), p.tasks.size() * sizeof(PanelTask), hipMemcpyHostToDevice));
    GP_HIP(hipMemcpy(d.wg_off.get(), p.wg_off.data(), p.wg_off.size() * sizeof(int), hipMemcpyHostToDevice));
    d.launches = p.launches;
    d.tmp_doubles = p.tmp_doubles;
    return panel_plans.emplace(key, std::move(d)).first->second;
}

// ---------------------------------------------------------------------------
// Cholesky: blocked right-looking with one panel of look-ahead.
//
// A panel of width Bo is factored in two parts. Its diagonal block (Bo x Bo)
// is factored in 64-wide steps and inverted on the `panel` stream; the rows
// below it are then one set of block products with that inverse on the
// `trail` stream, followed by a k-major copy of the panel that feeds the
// trailing update as the A^T operand. The latency-bound diagonal-block work
// is a few small launches per panel, so it slots in beside the update.
// Step k: the `panel` stream applies panel k to diagonal block k+1 (once
// step k-1's update has reached it) and factors that block, while the
// `trail` stream solves panel k's rows and applies panel k to everything
// below block k+1.
void cholesky_fp64(Context &c, double *A, long lda, int n_pad, double *Winv) { cholesky(c, A, lda, n_pad, Winv, false); }

int panel_width(int n_pad) { return outer_block(n_pad); }

void cholesky(Context &c, double *A, long lda, int n_pad, double *Winv, bool mixed, double *Wpanel, int extra)
{
    require(n_pad % 128 == 0 && lda >= n_pad && lda % 2 == 0, "cholesky: bad layout");
    require(extra >= 0 && extra % 64 == 0, "cholesky: extra rows must come in 64s");
    // rows of L and the rows below it that ride along (right-hand sides)
    const long rows = static_cast<long>(n_pad) + extra;
    const bool ahead = use_lookahead() && c.panel != c.main;
    hipStream_t T = ahead ? c.trail : c.main, P = ahead ? c.panel : c.main;
    // Two trailing workgroups per CU, minus the reserve. The FP32 product
    // would fit three per CU, but any cap above two per CU (552 to 744 on
    // 256 CUs) slowed the panel stream 2x and the n = 32768 mixed request
    // by 6 ms on MI355X.
    const int cap = ahead ? std::max(1, 2 * c.cus - trailing_reserve()) : 0;
    const int Bo = outer_block(n_pad);
    const int nblk = static_cast<int>(ceil_div(n_pad, Bo));
    k::set_int(c.main, c.info.get(), INT_MAX);
    // k-major copy of the factored panel's rows below its diagonal block
    // (FP32 for mixed precision), the diagonal block's inverse and scratch
    DevBuf<float> p32;
    DevBuf<double> p64;
    if (mixed)
        p32 = DevBuf<float>(c, static_cast<std::size_t>(Bo) * rows);
    else
        p64 = DevBuf<double>(c, static_cast<std::size_t>(Bo) * rows);
    // inverses of the panels' diagonal blocks: kept (Wpanel, block kb at
    // kb * Bo * Bo) or two scratch blocks used by alternate panels (the
    // panel stream factors panel k+2 only after the trailing stream's last
    // read of panel k's inverse)
    const int chain_wgs = ahead ? std::max(1, trailing_reserve()) : 4096, tail_wgs = ahead ? 64 : 4096;
    const int head_wgs = ahead ? head_workgroups() : 4096;
    DevBuf<double> Wscratch, tmp(c, static_cast<std::size_t>(std::max(1L, c.panel_plan(Bo / 64, chain_wgs, tail_wgs).tmp_doubles)));
    DevBuf<double> Xs(c, static_cast<std::size_t>(Bo) * rows);  // the panel's rows below, before the copy back
    DevBuf<double> Xh(c, static_cast<std::size_t>(Bo) * Bo);     // the next diagonal block's rows of the panel, solved
    DevBuf<int> queues(c, static_cast<std::size_t>(nblk));          // tile counters of the capped trailing updates, one per step
    if (!Wpanel)
        Wscratch = DevBuf<double>(c, static_cast<std::size_t>(2) * Bo * Bo);
    auto Wblock = [&](int p0)
    {
        const long kb = p0 / Bo;
        return Wpanel ? Wpanel + kb * Bo * Bo : Wscratch.get() + (kb % 2) * Bo * Bo;
    };

    std::vector<hipEvent_t> used;
    auto record = [&](hipStream_t s)
    {
        hipEvent_t e = c.event();
        used.push_back(e);
        GP_HIP(hipEventRecord(e, s));
        return e;
    };
    auto wait = [&](hipStream_t s, hipEvent_t e, hipStream_t from)
    {
        if (e && s != from)
            GP_HIP(hipStreamWaitEvent(s, e, 0));
    };
    k::PanelTaskArgs args{};
    args.A = A;
    args.lda = lda;
    args.Winv = Winv;
    args.info = c.info.get();
    // The diagonal block: its factor (in place), the inverses of its 64x64
    // blocks (Winv) and its whole inverse (Wblock), as the task schedule's
    // launches; the chain launches stay within the slots the capped
    // trailing update leaves free.
    auto diagonal = [&](int p0, int pe)
    {
        Timed t(c, P, "panel");
        k::PanelTaskArgs a = args;
        a.p0 = p0;
        a.nb = (pe - p0) / 64;
        a.Wb = Wblock(p0);
        a.ldw = Bo;
        a.tmp = tmp.get();
        const auto &plan = c.panel_plan(a.nb, chain_wgs, tail_wgs);
        for (const auto &l : plan.launches)
            k::panel_tasks(P, a, plan.tasks.get(), plan.wg_off.get() + l.first, l.groups);
    };
    // Panel [p0, pe)'s contribution to the next diagonal block [pe, q1),
    // on the panel stream: the block's rows of the panel solved with the
    // panel's inverse into Xh (then `solved` is recorded: the trailing
    // stream may overwrite those rows), and the block's update from them.
    auto head = [&](int p0, int pe, int q1, hipEvent_t &solved)
    {
        Timed t(c, P, "panel");
        k::PanelTaskArgs a = args;
        a.p0 = pe;
        a.nb = (q1 - pe) / 64;
        a.hp0 = p0;
        a.nbh = (pe - p0) / 64;
        a.Wh = Wblock(p0);
        a.ldwh = Bo;
        a.Xh = Xh.get();
        a.ldx = Bo;
        const auto &plan = c.head_plan(a.nbh, a.nb, head_wgs);
        require(plan.launches.size() == 2, "cholesky: unexpected head plan");
        k::panel_tasks(P, a, plan.tasks.get(), plan.wg_off.get() + plan.launches[0].first, plan.launches[0].groups);
        solved = record(P);
        k::panel_tasks(P, a, plan.tasks.get(), plan.wg_off.get() + plan.launches[1].first, plan.launches[1].groups);
    };
    // The panel's rows below its diagonal block: X = A W^T with W the
    // inverse of the diagonal block (k-major), one launch of 64x64 block
    // products into scratch (W is lower triangular, so column block c needs
    // c + 1 products; the launch pairs c with its mirror), then,
    // once the panel stream has read the next block's rows (`solved`), one
    // pass that writes the k-major copy for the trailing update and X back
    // into the panel.
    auto rows_below = [&](int p0, int pe, hipEvent_t solved)
    {
        Timed t(c, T, "panel_rows");
        const int w = pe - p0, R = static_cast<int>(rows - pe);
        double *panel = A + static_cast<long>(pe) * lda + p0;
        k::rows_solve(T, A, lda, pe, p0, R, w, Wblock(p0), Bo, Xs.get(), w);
        wait(T, solved, P);
        if (mixed)
            k::panel_to_f32t(T, Xs.get(), w, 0, 0, R, w, p32.get(), rows, panel, lda);
        else
            k::panel_to_f64t(T, Xs.get(), w, 0, 0, R, w, p64.get(), rows, panel, lda);
    };
    // Panel [p0, pe)'s update of everything below the next diagonal block:
    // rows q1 .., columns pe .. (the next panel's columns whole, the rest
    // lower), one capped launch.
    auto trailing = [&](int p0, int pe, int q1)
    {
        Timed t(c, T, "syrk_trailing");
        double *C = A + static_cast<long>(q1) * lda + pe;
        const int M = static_cast<int>(rows - q1), N = n_pad - pe, K = pe - p0;
        if (mixed)
            k::gemm_f32_trailing(T, p32.get(), rows, q1 - pe, 0, C, lda, M, N, K, q1 - pe, cap, queues.get() + p0 / Bo);
        else
            trailing_kmajor(T, p64.get(), rows, q1 - pe, 0, q1 - pe, C, lda, M, N, K, cap, queues.get() + p0 / Bo);
    };
    {
        Timed whole(c, c.main, "cholesky");
        wait(P, record(c.main), c.main);
        wait(T, record(c.main), c.main);
        GP_HIP(hipMemsetAsync(queues.get(), 0, sizeof(int) * nblk, T));
        diagonal(0, std::min(n_pad, Bo));
        // Step k: the panel stream prepares and factors block k+1 (its head
        // needs step k-1's trailing update); the trailing stream solves
        // panel k's rows (they need block k's inverse) and applies panel k
        // below block k+1.
        hipEvent_t trailed = nullptr;
        for (int kb = 0; kb + 1 < nblk; kb++)
        {
            const int p0 = kb * Bo, pe = std::min(n_pad, p0 + Bo), q1 = std::min(n_pad, pe + Bo);
            hipEvent_t factored = record(P), solved = nullptr;
            wait(P, trailed, T);
            head(p0, pe, q1, solved);
            diagonal(pe, q1);
            wait(T, factored, P);
            rows_below(p0, pe, solved);
            trailed = nullptr;
            if (q1 < rows)
            {
                trailing(p0, pe, q1);
                trailed = record(T);
            }
        }
        if (extra > 0)
        {
            // the rows below, against the last diagonal block
            wait(T, record(P), P);
            rows_below((nblk - 1) * Bo, n_pad, nullptr);
        }
        wait(c.main, record(T), T);
        wait(c.main, record(P), P);
    }
    c.sync();
    for (auto e : used)
        c.recycle(e);
    int info = INT_MAX;
    GP_HIP(hipMemcpy(&info, c.info.get(), sizeof(int), hipMemcpyDefault));
    if (info != INT_MAX)
        throw Error(GPRAT_ERR_NOT_SPD,
                    "covariance matrix is not positive definite (first non-positive pivot at column "
                        + std::to_string(info) + ")");
}

// alpha = K^-1 y by the blocked forward and backward solves. With the
// factorisation's wide diagonal-block inverses (Wpanel, k-major, wblock
// wide) each direction takes n_pad / wblock steps of two matrix-vector
// kernels (the diagonal block's inverse, then the coupling to the rest);
// otherwise 64-wide steps of one kernel each against Winv.
void solve_spd(Context &c, const double *L, long lda, const double *Winv, int n_pad, const double *y, double *work, double *out,
               const double *Wpanel, int wblock)
{
    hipStream_t s = c.main;
    Timed t(c, s, "solve");
    // work <- y (rhs of the forward solve, updated in place)
    GP_HIP(hipMemcpyAsync(work, y, sizeof(double) * n_pad, hipMemcpyDeviceToDevice, s));
    if (Wpanel)
    {
        require(wblock % 64 == 0 && wblock > 0, "solve_spd: bad block width");
        const int nblk = static_cast<int>(ceil_div(n_pad, wblock));
        auto WJ = [&](int J) { return Wpanel + static_cast<long>(J) * wblock * wblock; };
        for (int J = 0; J < nblk; J++)
        {
            const int j0 = J * wblock, nb = std::min(wblock, n_pad - j0), j1 = j0 + nb;
            k::coldot(s, WJ(J), wblock, nb, nb, work + j0, out + j0, false);
            if (j1 < n_pad)
                k::rowdot(s, L + static_cast<long>(j1) * lda + j0, lda, nb, n_pad - j1, out + j0, work + j1, true);
        }
        // out holds z = L^-1 y, updated in place; x = L^-T z goes to work
        for (int J = nblk - 1; J >= 0; J--)
        {
            const int j0 = J * wblock, nb = std::min(wblock, n_pad - j0);
            k::rowdot(s, WJ(J), wblock, nb, nb, out + j0, work + j0, false);
            if (j0 > 0)
                k::coldot(s, L + static_cast<long>(j0) * lda, lda, nb, j0, work + j0, out, true);
        }
    }
    else
    {
        const int nb = n_pad / 64;
        for (int j = 0; j < nb; j++)
            k::solve_lower_step(s, L, lda, Winv, work, out, n_pad, j);
        // out holds L^-1 y; use it as the rhs of the backward solve, result in work
        for (int j = nb - 1; j >= 0; j--)
            k::solve_upper_step(s, L, lda, Winv, out, work, n_pad, j);
    }
    GP_HIP(hipMemcpyAsync(out, work, sizeof(double) * n_pad, hipMemcpyDeviceToDevice, s));
}

// out = L^-1 B, blocked right-looking in NBK-wide steps: the diagonal step
// out_J = W_J B_J uses the k-major inverse of the diagonal block (the
// factorisation's panel inverses when given, NBK = their width, else
// 256-wide inverses built here), the update B[below] -= L[below, J] out_J
// reads a k-major copy of L's column panel, so both products take whole
// 1 KiB rows of their A operand. B is overwritten. With tri_rhs, B is lower
// triangular (B = I for L^-1) and each step's columns stop at the end of
// its block.
void trsm_blocked(Context &c, const double *L, long lda, const double *Winv, const double *Wpanel, int wblock, int n_pad, double *B,
                  long ldb, int ncols, bool tri_rhs, double *out, long ldo)
{
    hipStream_t s = c.main;
    const int NBK = Wpanel ? wblock : 256;
    require(NBK % 64 == 0 && NBK > 0, "trsm_blocked: bad block width");
    const int nblk = static_cast<int>(ceil_div(n_pad, NBK));
    DevBuf<double> Wt;
    if (!Wpanel)
    {
        Wt = DevBuf<double>(c, static_cast<std::size_t>(nblk) * NBK * NBK);
        k::invert_blocks256(s, L, lda, Winv, 0, n_pad, Wt.get(), NBK, static_cast<long>(NBK) * NBK);
    }
    const double *W = Wpanel ? Wpanel : Wt.get();
    DevBuf<double> Pt(c, static_cast<std::size_t>(NBK) * n_pad);
    for (int J = 0; J < nblk; J++)
    {
        const int j0 = J * NBK, nb = std::min(NBK, n_pad - j0), j1 = j0 + nb;
        const int cols = tri_rhs ? std::min(ncols, j1) : ncols;
        double *oJ = out + static_cast<long>(j0) * ldo;
        {
            k::GemmArgs g{};
            g.A = W + static_cast<long>(J) * NBK * NBK;
            g.lda = NBK;
            g.a_trans = true;
            g.B = B + static_cast<long>(j0) * ldb;
            g.ldb = ldb;
            g.C = oJ;
            g.ldc = ldo;
            g.M = nb;
            g.N = cols;
            g.K = nb;
            g.set = true;
            g.k_to_row = true;  // W_J is lower triangular
            g.role = k::GemmRole::LeafT;
            k::gemm_f64(s, g);
        }
        if (j1 >= n_pad)
            break;
        k::panel_to_f64t(s, L, lda, j1, j0, n_pad - j1, nb, Pt.get(), n_pad);
        k::GemmArgs u{};
        u.A = Pt.get();
        u.lda = n_pad;
        u.a_trans = true;
        u.B = oJ;
        u.ldb = ldo;
        u.C = B + static_cast<long>(j1) * ldb;
        u.ldc = ldb;
        u.M = n_pad - j1;
        u.N = cols;
        u.K = nb;
        u.role = k::GemmRole::SolveT;
        k::gemm_f64(s, u);
    }
}

// Precision-specific steps of the variance solve: the diagonal step
// V_J = W_J B_J (W_J k-major), the update B -= P V_J from the k-major copy
// P of L's column panel, the squares of V_J, and the copy itself.
struct VarianceF64
{
    using T = double;
    static void leaf(hipStream_t s, const T *W, int ldw, const T *B, long ldb, T *V, int nb, int cols)
    {
        if (nb % 64 == 0 && cols % 64 == 0)
        {
            k::leaf_solve(s, W, ldw, nb, B, ldb, cols, V, ldb);
            return;
        }
        k::GemmArgs g{};
        g.A = W;
        g.lda = ldw;
        g.a_trans = true;
        g.B = B;
        g.ldb = ldb;
        g.C = V;
        g.ldc = ldb;
        g.M = nb;
        g.N = cols;
        g.K = nb;
        g.set = true;
        g.k_to_row = true;  // W_J is lower triangular
        g.role = k::GemmRole::LeafT;
        k::gemm_f64(s, g);
    }
    static void update(hipStream_t s, const T *P, long ldp, const T *V, long ldv, T *B, long ldb, int M, int cols, int K)
    {
        k::GemmArgs u{};
        u.A = P;
        u.lda = ldp;
        u.a_trans = true;
        u.B = V;
        u.ldb = ldv;
        u.C = B;
        u.ldc = ldb;
        u.M = M;
        u.N = cols;
        u.K = K;
        u.role = k::GemmRole::SolveT;
        k::gemm_f64(s, u);
    }
    static void squares(hipStream_t s, const T *V, long ldv, int rows, int mt, double *colsq) { k::colsq_accumulate(s, V, ldv, rows, mt, colsq); }
    static void panel(hipStream_t s, const double *L, long lda, long j1, long j0, int R, int W, T *P, long ldp)
    {
        k::panel_to_f64t(s, L, lda, j1, j0, R, W, P, ldp);
    }
};

struct VarianceF32
{
    using T = float;
    static void leaf(hipStream_t s, const T *W, int ldw, const T *B, long ldb, T *V, int nb, int cols)
    {
        k::gemm_f32_tn(s, W, ldw, B, ldb, V, ldb, nb, cols, nb, true, 0, true);  // W_J lower triangular
    }
    static void update(hipStream_t s, const T *P, long ldp, const T *V, long ldv, T *B, long ldb, int M, int cols, int K)
    {
        k::gemm_f32_tn(s, P, ldp, V, ldv, B, ldb, M, cols, K, false);
    }
    static void squares(hipStream_t s, const T *V, long ldv, int rows, int mt, double *colsq)
    {
        k::colsq_accumulate_f32(s, V, ldv, rows, mt, colsq);
    }
    static void panel(hipStream_t s, const double *L, long lda, long j1, long j0, int R, int W, T *P, long ldp)
    {
        k::panel_to_f32t(s, L, lda, j1, j0, R, W, P, ldp);
    }
};

// V = L^-1 B (B n_pad x m_pad, overwritten), blocked right-looking in
// NBK-row steps with the k-major block inverses W: V_J = W_J B_J goes to
// consume(stream, V_J, rows) (V_J is NBK x m_pad, leading dimension m_pad,
// valid until the next step), then B[below] -= P_J V_J with P_J the k-major
// copy of L's column panel J. V itself is never stored whole.
template <class Ops, class Consume>
static void blocked_solve(Context &c, const double *L, long lda, const typename Ops::T *W, int NBK, int n_pad, typename Ops::T *B, int m_pad,
                          Consume consume)
{
    using T = typename Ops::T;
    hipStream_t s = c.main;
    // the k-major copies of L's panels are made on the panel stream, one
    // step ahead into alternating buffers, so they overlap the diagonal
    // steps (which fill half the device) instead of preceding each update
    hipStream_t side = c.panel;
    const int nblk = static_cast<int>(ceil_div(n_pad, NBK));
    DevBuf<T> V(c, static_cast<std::size_t>(NBK) * m_pad);
    DevBuf<T> P[2] = { DevBuf<T>(c, static_cast<std::size_t>(NBK) * n_pad), DevBuf<T>(c, static_cast<std::size_t>(NBK) * n_pad) };
    std::vector<hipEvent_t> used;
    auto record = [&](hipStream_t st)
    {
        hipEvent_t e = c.event();
        used.push_back(e);
        GP_HIP(hipEventRecord(e, st));
        return e;
    };
    hipEvent_t freed[2] = { record(s), nullptr };  // the buffer's last update is done (side waits)
    freed[1] = freed[0];
    auto copy = [&](int J)
    {
        const int j0 = J * NBK, nb = std::min(NBK, n_pad - j0), j1 = j0 + nb;
        if (side != s)
            GP_HIP(hipStreamWaitEvent(side, freed[J % 2], 0));
        Ops::panel(side, L, lda, j1, j0, n_pad - j1, nb, P[J % 2].get(), n_pad);
        return record(side);
    };
    hipEvent_t ready = (NBK < n_pad) ? copy(0) : nullptr;
    for (int J = 0; J < nblk; J++)
    {
        const int j0 = J * NBK, nb = std::min(NBK, n_pad - j0), j1 = j0 + nb;
        hipEvent_t next = nullptr;
        if (j1 + NBK < n_pad)  // step J+1 has rows below its block
            next = copy(J + 1);
        Ops::leaf(s, W + static_cast<long>(J) * NBK * NBK, NBK, B + static_cast<long>(j0) * m_pad, m_pad, V.get(), nb, m_pad);
        consume(s, static_cast<const T *>(V.get()), nb);
        if (j1 >= n_pad)
            break;
        if (side != s)
            GP_HIP(hipStreamWaitEvent(s, ready, 0));
        {
            Timed t(c, s, "trsm_update");
            Ops::update(s, P[J % 2].get(), n_pad, V.get(), m_pad, B + static_cast<long>(j1) * m_pad, m_pad, n_pad - j1, m_pad, nb);
        }
        freed[J % 2] = record(s);
        ready = next;
    }
    if (side != s)
        GP_HIP(hipStreamWaitEvent(s, record(side), 0));
    c.sync();
    for (auto e : used)
        c.recycle(e);
}

// colsq[j] = sum_i V[i][j]^2 over the solve above, for j < mt; squares are
// summed in step order.
template <class Ops>
static void variance_solve(Context &c, const double *L, long lda, const typename Ops::T *W, int NBK, int n_pad, typename Ops::T *B,
                           int m_pad, int mt, double *colsq)
{
    k::fill(c.main, colsq, m_pad, 0.0);
    blocked_solve<Ops>(c, L, lda, W, NBK, n_pad, B, m_pad,
                       [&](hipStream_t s, const typename Ops::T *V, int rows) { Ops::squares(s, V, m_pad, rows, mt, colsq); });
}

// Column sums of squares of V = L^-1 K*^T for a mixed-precision model: the
// same 256-step blocked solve as trsm_blocked, with FP32 copies of the block
// inverses, of L's column panels and of K*^T on FP32 matrix cores; the
// squares are summed in FP64.
static void variance_colsq_f32(Context &c, const FactorView &m, const double *z, int mt, int m_pad, double *colsq)
{
    hipStream_t s = c.main;
    const int NBK = m.Wpanel ? m.wblock : 256;
    const int n_pad = m.n_pad;
    const int nblk = static_cast<int>(ceil_div(n_pad, NBK));
    const std::size_t wcount = static_cast<std::size_t>(nblk) * NBK * NBK;
    DevBuf<float> Wt32(c, wcount);
    if (m.Wpanel)
        k::to_f32(s, m.Wpanel, Wt32.get(), static_cast<long>(wcount));
    else
    {
        DevBuf<double> Wt(c, wcount);
        k::invert_blocks256(s, m.L, m.lda, m.Winv, 0, n_pad, Wt.get(), NBK, static_cast<long>(NBK) * NBK);
        k::to_f32(s, Wt.get(), Wt32.get(), static_cast<long>(wcount));
        c.sync();  // Wt goes back to the pool
    }
    DevBuf<float> B(c, static_cast<std::size_t>(n_pad) * m_pad);
    {
        Timed t(c, s, "assemble");
        k::assemble_cross_t_f32(s, m.x, z, m.r, m.sek, B.get(), m_pad, m.n, n_pad, mt, m_pad);
    }
    Timed t(c, s, "trsm");
    variance_solve<VarianceF32>(c, m.L, m.lda, Wt32.get(), NBK, n_pad, B.get(), m_pad, mt, colsq);
}

// X = L^-T V in FP32 (V n_pad x m_pad in B, overwritten by X), blocked in
// NBK-row steps from the last block up: X_J = W_J^T V_J (WT: the transposed
// block inverses, k-major, so the product's A operand loads whole rows),
// then V[above] -= L[J, above]^T X_J from an FP32 copy of L's row block J
// (rows j0 .. j0 + nb, columns 0 .. j0: the k-major operand layout as it
// stands). The row-block copies are made on the panel stream one step
// ahead into alternating buffers.
static void backward_solve_f32(Context &c, const double *L, long lda, const float *WT, int NBK, int n_pad, float *B, int m_pad)
{
    hipStream_t s = c.main;
    hipStream_t side = c.panel;
    const int nblk = static_cast<int>(ceil_div(n_pad, NBK));
    DevBuf<float> T(c, static_cast<std::size_t>(NBK) * m_pad);
    DevBuf<float> R[2] = { DevBuf<float>(c, static_cast<std::size_t>(NBK) * n_pad), DevBuf<float>(c, static_cast<std::size_t>(NBK) * n_pad) };
    std::vector<hipEvent_t> used;
    auto record = [&](hipStream_t st)
    {
        hipEvent_t e = c.event();
        used.push_back(e);
        GP_HIP(hipEventRecord(e, st));
        return e;
    };
    hipEvent_t freed[2] = { record(s), nullptr };
    freed[1] = freed[0];
    auto copy = [&](int J)
    {
        const int j0 = J * NBK, nb = std::min(NBK, n_pad - j0);
        if (side != s)
            GP_HIP(hipStreamWaitEvent(side, freed[J % 2], 0));
        k::rows_to_f32(side, L, lda, j0, nb, j0, R[J % 2].get(), n_pad);
        return record(side);
    };
    hipEvent_t ready = nblk > 1 ? copy(nblk - 1) : nullptr;
    for (int J = nblk - 1; J >= 0; J--)
    {
        const int j0 = J * NBK, nb = std::min(NBK, n_pad - j0);
        hipEvent_t next = J - 1 >= 1 ? copy(J - 1) : nullptr;
        float *BJ = B + static_cast<long>(j0) * m_pad;
        k::gemm_f32_tn(s, WT + static_cast<long>(J) * NBK * NBK, NBK, BJ, m_pad, T.get(), m_pad, nb, m_pad, nb, true);
        GP_HIP(hipMemcpyAsync(BJ, T.get(), sizeof(float) * nb * m_pad, hipMemcpyDeviceToDevice, s));
        if (j0 == 0)
            break;
        if (side != s)
            GP_HIP(hipStreamWaitEvent(s, ready, 0));
        {
            Timed t(c, s, "trsm_update");
            k::gemm_f32_tn(s, R[J % 2].get(), n_pad, T.get(), m_pad, B, m_pad, j0, m_pad, nb, false);
        }
        freed[J % 2] = record(s);
        ready = next;
    }
    if (side != s)
        GP_HIP(hipStreamWaitEvent(s, record(side), 0));
    c.sync();
    for (auto e : used)
        c.recycle(e);
}

// Latent variance of a mixed-precision model, refined in FP64 (k_refine.hip):
// X~ = K~^-1 K*^T by the FP32 blocked forward solve (its column squares are
// the unrefined q0 = |L~^-1 k|^2) and the FP32 backward solve, on matrix
// cores with FP32 copies of the block inverses; then, in FP64, the residual
// product Z = K_sl X~ and per column q1 = 2 k^T x~ - (2 x~^T z + d |x~|^2),
// var = v - q1. The result's error is e^T K e for the solves' error e, so
// it is second order in the FP32 and factor errors. Returns max |q1 - q0| / v.
static double variance_refined(Context &c, const FactorView &m, const double *z, int mt, int m_pad, double *var)
{
    hipStream_t s = c.main;
    const int NBK = m.Wpanel ? m.wblock : 256;
    const int n_pad = m.n_pad;
    const int nblk = static_cast<int>(ceil_div(n_pad, NBK));
    const std::size_t wcount = static_cast<std::size_t>(nblk) * NBK * NBK;
    const std::size_t elems = static_cast<std::size_t>(n_pad) * m_pad;
    DevBuf<float> W32(c, wcount), WT32(c, wcount);
    DevBuf<double> Wt, Ktmp;
    const double *W64 = m.Wpanel;
    if (!W64)
    {
        Wt = DevBuf<double>(c, wcount);
        k::invert_blocks256(s, m.L, m.lda, m.Winv, 0, n_pad, Wt.get(), NBK, static_cast<long>(NBK) * NBK);
        W64 = Wt.get();
    }
    k::to_f32(s, W64, W32.get(), static_cast<long>(wcount));
    k::transpose_blocks_f32(s, W64, WT32.get(), NBK, nblk);
    const double *Kup = m.Kup;
    long ldk = m.ldk;
    if (!Kup)
    {
        Ktmp = DevBuf<double>(c, static_cast<std::size_t>(n_pad) * n_pad);
        k::assemble_upper(s, m.x, m.r, m.sek, Ktmp.get(), n_pad, m.n, n_pad);
        Kup = Ktmp.get();
        ldk = n_pad;
    }
    DevBuf<float> B32(c, elems);
    DevBuf<double> q0(c, m_pad);
    {
        Timed t(c, s, "assemble");
        k::assemble_cross_t_f32(s, m.x, z, m.r, m.sek, B32.get(), m_pad, m.n, n_pad, mt, m_pad);
    }
    {
        Timed t(c, s, "trsm");
        k::fill(s, q0.get(), m_pad, 0.0);
        int J = 0;
        float *b32 = B32.get();
        blocked_solve<VarianceF32>(c, m.L, m.lda, W32.get(), NBK, n_pad, b32, m_pad,
                                   [&](hipStream_t st, const float *V, int rows)
                                   {
                                       k::colsq_accumulate_f32(st, V, m_pad, rows, mt, q0.get());
                                       // B_J has been consumed by the diagonal step: V_J takes its place
                                       GP_HIP(hipMemcpyAsync(b32 + static_cast<long>(J) * NBK * m_pad, V, sizeof(float) * rows * m_pad,
                                                             hipMemcpyDeviceToDevice, st));
                                       J++;
                                   });
        backward_solve_f32(c, m.L, m.lda, WT32.get(), NBK, n_pad, b32, m_pad);
    }
    DevBuf<double> X64(c, elems), B64(c, elems), Z(c, elems);
    k::to_f64(s, B32.get(), X64.get(), static_cast<long>(elems));
    {
        Timed t(c, s, "assemble");
        k::assemble_cross_t(s, m.x, z, m.r, m.sek, B64.get(), m_pad, m.n, n_pad, mt, m_pad);
    }
    const int splits = k::quad_form_splits(n_pad, m_pad);
    DevBuf<double> part(c, static_cast<std::size_t>(splits) * 3 * m_pad), corr(c, m_pad);
    {
        Timed t(c, s, "residual");
        k::GemmArgs g{};
        g.A = Kup;  // A(i, k) = Kup[k][i] = K(i, k) for k < i
        g.lda = ldk;
        g.a_trans = true;
        g.B = X64.get();
        g.ldb = m_pad;
        g.C = Z.get();
        g.ldc = m_pad;
        g.M = n_pad;
        g.N = m_pad;
        g.K = n_pad;
        g.set = true;
        g.k_before_tile = true;
        g.role = k::GemmRole::Residual;
        k::gemm_f64(s, g);
    }
    {
        Timed t(c, s, "refine");
        k::diag_strict_product(s, Kup, ldk, X64.get(), m_pad, Z.get(), m_pad, n_pad, m_pad);
        k::quad_form(s, B64.get(), X64.get(), Z.get(), m_pad, n_pad, m_pad, m.sek.vertical + m.sek.noise, m.sek.vertical, q0.get(), mt,
                     part.get(), splits, var, corr.get());
    }
    std::vector<double> hc(mt);
    GP_HIP(hipMemcpyAsync(hc.data(), corr.get(), sizeof(double) * mt, hipMemcpyDefault, s));
    c.sync();
    double worst = 0.0;
    for (double v : hc)
        worst = std::max(worst, std::fabs(v));
    return m.sek.vertical != 0.0 ? worst / std::fabs(m.sek.vertical) : worst;
}

// ---------------------------------------------------------------------------
// Model operations

static void upload(Context &c, double *dst, const double *src, std::size_t count)
{
    if (count)
        GP_HIP(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDefault, c.main));
}

void factor(Context &c, Model &m, const double *x, std::size_t x_len, int n, int r, const double theta[3], int precision)
{
    require(n >= 1, "need at least one training point");
    require(r >= 1, "n_reg must be >= 1");
    require(x != nullptr && x_len >= static_cast<std::size_t>(n) + r - 1, "training input must hold n + n_reg - 1 values");
    require(theta != nullptr, "theta is null");
    require(precision == GPRAT_PRECISION_FP64 || precision == GPRAT_PRECISION_MIXED, "unknown precision mode");
    c.make_current();
    m.ctx = &c;
    m.n = n;
    m.r = r;
    m.n_pad = static_cast<int>(padded(n));
    m.lda = leading_dim(m.n_pad);
    m.sek = SEK{ theta[0], theta[1], theta[2] };
    m.precision = precision;
    const std::size_t xl = static_cast<std::size_t>(n) + r - 1;
    m.x = DevBuf<double>(c, xl);
    m.L = DevBuf<double>(c, static_cast<std::size_t>(m.n_pad) * m.lda);
    m.Winv = DevBuf<double>(c, static_cast<std::size_t>(m.n_pad) * 64);
    m.wblock = panel_width(m.n_pad);
    m.Wpanel = DevBuf<double>(c, static_cast<std::size_t>(ceil_div(m.n_pad, m.wblock)) * m.wblock * m.wblock);
    upload(c, m.x.get(), x, xl);
    {
        Timed t(c, c.main, "assemble");
        k::assemble_lower(c.main, m.x.get(), r, m.sek, m.L.get(), m.lda, n, m.n_pad);
    }
    cholesky(c, m.L.get(), m.lda, m.n_pad, m.Winv.get(), precision == GPRAT_PRECISION_MIXED, m.Wpanel.get());
    if (precision == GPRAT_PRECISION_MIXED)
    {
        // the covariance's strictly lower part, for the FP64 residual of the
        // variance refinement, goes to the factor's free upper triangle
        Timed t(c, c.main, "assemble");
        k::assemble_upper(c.main, m.x.get(), r, m.sek, m.L.get(), m.lda, n, m.n_pad);
    }
}

// FP64 iterative refinement of alpha against a factor of reduced accuracy:
// r = y - K alpha with K regenerated from the series in FP64, delta from the
// factor, alpha += delta, until the correction, or the next one as the
// contraction so far predicts it, is at most 1e-12 max|alpha|, or the
// corrections stop shrinking.
void refine_alpha(Model &m, double *work)
{
    Context &c = *m.ctx;
    hipStream_t s = c.main;
    const int n = m.n;
    const int splits = k::mean_splits(n, n);
    DevBuf<double> partial(c, static_cast<std::size_t>(splits) * n);
    DevBuf<double> kx(c, m.n_pad), r(c, m.n_pad), delta(c, m.n_pad);
    k::fill(s, r.get(), m.n_pad, 0.0);
    std::vector<double> hd(n), ha(n);
    double prev = 1e300;
    m.refine_iterations = 0;
    for (int it = 0; it < 30; it++)
    {
        {
            Timed t(c, s, "refine");
            k::predict_mean(s, m.x.get(), m.x.get(), m.alpha.get(), m.r, m.sek, n, n, partial.get(), splits, kx.get());
            k::residual_finish(s, m.y.get(), kx.get(), m.alpha.get(), m.sek.noise, n, r.get());
        }
        solve_spd(c, m.L.get(), m.lda, m.Winv.get(), m.n_pad, r.get(), work, delta.get(), m.Wpanel.get(), m.wblock);
        GP_HIP(hipMemcpyAsync(hd.data(), delta.get(), sizeof(double) * n, hipMemcpyDefault, s));
        GP_HIP(hipMemcpyAsync(ha.data(), m.alpha.get(), sizeof(double) * n, hipMemcpyDefault, s));
        c.sync();
        double dmax = 0.0, amax = 0.0;
        for (int i = 0; i < n; i++)
        {
            dmax = std::max(dmax, std::fabs(hd[i]));
            amax = std::max(amax, std::fabs(ha[i]));
        }
        // a correction that did not shrink is not applied: alpha keeps the
        // best iterate so far
        if (it > 0 && dmax >= prev)
            break;
        k::add_vec(s, m.alpha.get(), delta.get(), n);
        m.refine_iterations = it + 1;
        m.refine_last = amax > 0.0 ? dmax / amax : 0.0;
        // the next correction is about (dmax / prev) * dmax: when that is
        // already below the target, this step has reached it
        const bool converged = dmax <= 1e-12 * amax || (it > 0 && dmax * (dmax / prev) <= 1e-12 * amax);
        if (!converged && dmax < 0.5 * prev)
        {
            prev = dmax;
            continue;
        }
        break;
    }
}

void attach_targets(Model &m, const double *y, std::size_t y_len)
{
    Context &c = *m.ctx;
    require(y != nullptr && y_len >= static_cast<std::size_t>(m.n), "training output must hold n values");
    c.make_current();
    m.y = DevBuf<double>(c, m.n_pad);
    m.alpha = DevBuf<double>(c, m.n_pad);
    DevBuf<double> work(c, m.n_pad);
    k::fill(c.main, m.y.get(), m.n_pad, 0.0);
    upload(c, m.y.get(), y, static_cast<std::size_t>(m.n));
    solve_spd(c, m.L.get(), m.lda, m.Winv.get(), m.n_pad, m.y.get(), work.get(), m.alpha.get(), m.Wpanel.get(), m.wblock);
    if (m.precision == GPRAT_PRECISION_MIXED)
        refine_alpha(m, work.get());
    c.sync();
}

void fit(Context &c,
         Model &m,
         const double *x,
         std::size_t x_len,
         const double *y,
         std::size_t y_len,
         int n,
         int r,
         const double theta[3],
         int precision)
{
    require(y != nullptr && y_len >= static_cast<std::size_t>(n), "training output must hold n values");
    factor(c, m, x, x_len, n, r, theta, precision);
    attach_targets(m, y, y_len);
}

void predict_view(const FactorView &m, const double *z, std::size_t z_len, int mt, double *mean, double *var)
{
    Context &c = *m.ctx;
    require(mt >= 0, "predict: negative test count");
    if (mt == 0)
        return;
    require(z != nullptr && z_len >= static_cast<std::size_t>(mt) + m.r - 1,
            "predict: test input must hold m + n_reg - 1 values");
    require(mean != nullptr, "predict: mean output is null");
    c.make_current();
    const std::size_t zl = static_cast<std::size_t>(mt) + m.r - 1;
    DevBuf<double> dz(c, zl);
    upload(c, dz.get(), z, zl);
    const int splits = k::mean_splits(m.n, mt);
    DevBuf<double> partial(c, static_cast<std::size_t>(splits) * mt);
    DevBuf<double> dmean(c, mt);
    {
        Timed t(c, c.main, "mean");
        k::predict_mean(c.main, m.x, dz.get(), m.alpha, m.r, m.sek, m.n, mt, partial.get(), splits, dmean.get());
    }
    DevBuf<double> dvar;
    if (var && m.mixed && m.refine_variance)
    {
        const int m_pad = static_cast<int>(round_up(mt, 64));
        dvar = DevBuf<double>(c, mt);
        const double corr = variance_refined(c, m, dz.get(), mt, m_pad, dvar.get());
        if (m.variance_correction)
            *m.variance_correction = corr;
        GP_HIP(hipMemcpyAsync(var, dvar.get(), sizeof(double) * mt, hipMemcpyDefault, c.main));
    }
    else if (var && m.mixed)
    {
        const int m_pad = static_cast<int>(round_up(mt, 64));
        DevBuf<double> colsq(c, m_pad);
        dvar = DevBuf<double>(c, mt);
        variance_colsq_f32(c, m, dz.get(), mt, m_pad, colsq.get());
        k::variance_finish(c.main, colsq.get(), m.sek.vertical, mt, dvar.get());
        GP_HIP(hipMemcpyAsync(var, dvar.get(), sizeof(double) * mt, hipMemcpyDefault, c.main));
    }
    else if (var)
    {
        const int m_pad = static_cast<int>(round_up(mt, 64));
        const long ldb = m_pad;
        DevBuf<double> B(c, static_cast<std::size_t>(m.n_pad) * ldb);
        DevBuf<double> colsq(c, m_pad);
        dvar = DevBuf<double>(c, mt);
        {
            Timed t(c, c.main, "assemble");
            k::assemble_cross_t(c.main, m.x, dz.get(), m.r, m.sek, B.get(), ldb, m.n, m.n_pad, mt, m_pad);
        }
        {
            Timed t(c, c.main, "trsm");
            // the factorisation's block inverses, or 256-wide ones built here
            const int NBK = m.Wpanel ? m.wblock : 256;
            DevBuf<double> Wt;
            if (!m.Wpanel)
            {
                Wt = DevBuf<double>(c, static_cast<std::size_t>(ceil_div(m.n_pad, NBK)) * NBK * NBK);
                k::invert_blocks256(c.main, m.L, m.lda, m.Winv, 0, m.n_pad, Wt.get(), NBK, static_cast<long>(NBK) * NBK);
            }
            variance_solve<VarianceF64>(c, m.L, m.lda, m.Wpanel ? m.Wpanel : Wt.get(), NBK, m.n_pad, B.get(), m_pad, mt, colsq.get());
        }
        k::variance_finish(c.main, colsq.get(), m.sek.vertical, mt, dvar.get());
        GP_HIP(hipMemcpyAsync(var, dvar.get(), sizeof(double) * mt, hipMemcpyDefault, c.main));
    }
    GP_HIP(hipMemcpyAsync(mean, dmean.get(), sizeof(double) * mt, hipMemcpyDefault, c.main));
    c.sync();
}

// Mean and the whole latent posterior covariance Sigma = K** - V^T V,
// V = L^-1 K*^T, into cov (mt x mt, row-major). The reference forms Sigma
// tile by tile (gen_tile_full_prior_covariance, gp_algorithms.cpp:64-88;
// symmetric_matrix_matrix_tiled, tiled_algorithms.cpp:230-254) and returns
// only its diagonal (gp_functions.cpp:663-695); here Sigma starts as K**
// and every block V_J of the blocked solve subtracts V_J^T V_J on the
// matrix cores, in step order. FP64 throughout (a mixed-precision model's
// factor carries its own accuracy into Sigma).
void predict_full_cov_view(const FactorView &m, const double *z, std::size_t z_len, int mt, double *mean, double *cov)
{
    Context &c = *m.ctx;
    require(mt >= 0, "predict_full_cov: negative test count");
    if (mt == 0)
        return;
    require(cov != nullptr, "predict_full_cov: covariance output is null");
    predict_view(m, z, z_len, mt, mean, nullptr);
    c.make_current();
    const std::size_t zl = static_cast<std::size_t>(mt) + m.r - 1;
    DevBuf<double> dz(c, zl);
    upload(c, dz.get(), z, zl);
    const int m_pad = static_cast<int>(round_up(mt, 64));
    DevBuf<double> B(c, static_cast<std::size_t>(m.n_pad) * m_pad), S(c, static_cast<std::size_t>(m_pad) * m_pad);
    {
        Timed t(c, c.main, "assemble");
        k::assemble_cross_t(c.main, m.x, dz.get(), m.r, m.sek, B.get(), m_pad, m.n, m.n_pad, mt, m_pad);
        // K** (no noise term: the latent covariance), zero in the padding
        k::assemble_cross_t(c.main, dz.get(), dz.get(), m.r, m.sek, S.get(), m_pad, mt, m_pad, mt, m_pad);
    }
    {
        Timed t(c, c.main, "trsm");
        const int NBK = m.Wpanel ? m.wblock : 256;
        DevBuf<double> Wt;
        if (!m.Wpanel)
        {
            Wt = DevBuf<double>(c, static_cast<std::size_t>(ceil_div(m.n_pad, NBK)) * NBK * NBK);
            k::invert_blocks256(c.main, m.L, m.lda, m.Winv, 0, m.n_pad, Wt.get(), NBK, static_cast<long>(NBK) * NBK);
        }
        blocked_solve<VarianceF64>(c, m.L, m.lda, m.Wpanel ? m.Wpanel : Wt.get(), NBK, m.n_pad, B.get(), m_pad,
                                   [&](hipStream_t s, const double *V, int rows)
                                   {
                                       // Sigma -= V_J^T V_J
                                       k::GemmArgs g{};
                                       g.A = V;  // A(i, k) = V_J[k][i]
                                       g.lda = m_pad;
                                       g.a_trans = true;
                                       g.B = V;  // B(k, j) = V_J[k][j]
                                       g.ldb = m_pad;
                                       g.C = S.get();
                                       g.ldc = m_pad;
                                       g.M = m_pad;
                                       g.N = m_pad;
                                       g.K = rows;
                                       g.role = k::GemmRole::SolveT;
                                       k::gemm_f64(s, g);
                                   });
    }
    GP_HIP(hipMemcpy2DAsync(cov, sizeof(double) * mt, S.get(), sizeof(double) * m_pad, sizeof(double) * mt, mt, hipMemcpyDefault, c.main));
    c.sync();
}

void predict_full_cov(Model &m, const double *z, std::size_t z_len, int mt, double *mean, double *cov)
{
    predict_full_cov_view(FactorView{ m.ctx, m.x.get(), m.n, m.n_pad, m.r, m.lda, m.sek, m.L.get(), m.Winv.get(), m.alpha.get(),
                                      m.precision == GPRAT_PRECISION_MIXED, m.Wpanel.get(), m.wblock },
                          z, z_len, mt, mean, cov);
}

// Predictive mean through the fixed device scratch (no device allocation):
// the test series is staged at the scratch's start, then the mean's partial
// sums and the mean itself, each 256-byte aligned. Every partial sum is
// written before it is read, so nothing needs clearing. Calls on one GPU
// share its scratch: they are serialised by the context lock of the model,
// and calls on different contexts of one GPU must not overlap.
void predict_mean_noalloc(Model &m, const double *z, std::size_t z_len, int mt, double *mean)
{
    Context &c = *m.ctx;
    require(mt >= 1, "predict_mean_noalloc: need at least one test point");
    require(z != nullptr && z_len >= static_cast<std::size_t>(mt) + m.r - 1,
            "predict_mean_noalloc: test input must hold m + n_reg - 1 values");
    require(mean != nullptr, "predict_mean_noalloc: mean output is null");
    const std::size_t zl = static_cast<std::size_t>(mt) + m.r - 1;
    const int splits = k::mean_splits(m.n, mt);
    auto aligned = [](std::size_t bytes) { return (bytes + 255) / 256 * 256; };
    const std::size_t off_partial = aligned(zl * sizeof(double));
    const std::size_t off_mean = off_partial + aligned(static_cast<std::size_t>(splits) * mt * sizeof(double));
    const std::size_t need = off_mean + aligned(static_cast<std::size_t>(mt) * sizeof(double));
    if (need > k::noalloc_scratch_bytes())
        throw Error(GPRAT_ERR_ARGUMENT, "predict_mean_noalloc: " + std::to_string(mt) + " test points need " + std::to_string(need)
                                            + " bytes of scratch, more than the fixed " + std::to_string(k::noalloc_scratch_bytes()));
    c.make_current();
    char *base = static_cast<char *>(k::noalloc_scratch(c.device));
    double *dz = reinterpret_cast<double *>(base);
    double *partial = reinterpret_cast<double *>(base + off_partial);
    double *dmean = reinterpret_cast<double *>(base + off_mean);
    GP_HIP(hipMemcpyAsync(dz, z, zl * sizeof(double), hipMemcpyDefault, c.main));
    k::predict_mean(c.main, m.x.get(), dz, m.alpha.get(), m.r, m.sek, m.n, mt, partial, splits, dmean);
    GP_HIP(hipMemcpyAsync(mean, dmean, sizeof(double) * mt, hipMemcpyDefault, c.main));
    GP_HIP(hipStreamSynchronize(c.main));
}

// The rows of K* below K in one matrix, factored with it: on return they
// hold V^T = K* L^-T (the trailing updates and panel-row solves of every
// step carry the variance's forward solve, and fill the steps where the
// diagonal-block chain leaves the device idle), and var = v - their sums of
// squares. The results agree with fit() and predict() to rounding, not bit
// for bit: the test rows enlarge the late steps' trailing updates past the
// workgroup cap, whose split K halves then round some of L's elements
// differently, and the variance sums its squares in another order.
void predict_with_uncertainty_once(Context &c, const double *x, std::size_t x_len, const double *y, std::size_t y_len, int n, int r,
                                   const double theta[3], const double *z, std::size_t z_len, int mt, double *mean, double *var)
{
    require(n >= 1 && r >= 1 && mt >= 1, "predict: need training and test points");
    require(x != nullptr && x_len >= static_cast<std::size_t>(n) + r - 1, "training input must hold n + n_reg - 1 values");
    require(z != nullptr && z_len >= static_cast<std::size_t>(mt) + r - 1, "predict: test input must hold m + n_reg - 1 values");
    require(theta != nullptr && mean != nullptr && var != nullptr, "predict: null argument");
    require(y != nullptr && y_len >= static_cast<std::size_t>(n), "training output must hold n values");
    c.make_current();
    Model m;
    m.ctx = &c;
    m.n = n;
    m.r = r;
    m.n_pad = static_cast<int>(padded(n));
    m.lda = leading_dim(m.n_pad);
    m.sek = SEK{ theta[0], theta[1], theta[2] };
    m.precision = GPRAT_PRECISION_FP64;
    const int m_pad = static_cast<int>(round_up(mt, 64));
    // below K: the m test rows of K*, then a 64-row block whose first row is y
    const int extra = m_pad + 64;
    const std::size_t xl = static_cast<std::size_t>(n) + r - 1, zl = static_cast<std::size_t>(mt) + r - 1;
    m.x = DevBuf<double>(c, xl);
    m.L = DevBuf<double>(c, (static_cast<std::size_t>(m.n_pad) + extra) * m.lda);
    m.Winv = DevBuf<double>(c, static_cast<std::size_t>(m.n_pad) * 64);
    m.wblock = panel_width(m.n_pad);
    m.Wpanel = DevBuf<double>(c, static_cast<std::size_t>(ceil_div(m.n_pad, m.wblock)) * m.wblock * m.wblock);
    DevBuf<double> dz(c, zl);
    upload(c, m.x.get(), x, xl);
    upload(c, dz.get(), z, zl);
    double *rows = m.L.get() + static_cast<long>(m.n_pad) * m.lda;
    double *yrow = rows + static_cast<long>(m_pad) * m.lda;
    {
        Timed t(c, c.main, "assemble");
        k::assemble_lower(c.main, m.x.get(), r, m.sek, m.L.get(), m.lda, n, m.n_pad);
        k::assemble_cross_rows(c.main, m.x.get(), dz.get(), r, m.sek, rows, m.lda, n, m.n_pad, mt, m_pad);
        k::fill(c.main, yrow, 64L * m.lda, 0.0);
        GP_HIP(hipMemcpyAsync(yrow, y, sizeof(double) * n, hipMemcpyDefault, c.main));
    }
    // the rows below K come out as K* L^-T (test row t: v_t = L^-1 k_t) and
    // (L^-1 y)^T, so the mean is K* K^-1 y = V (L^-1 y) and the variance
    // v - |v_t|^2: no alpha and no separate solve
    cholesky(c, m.L.get(), m.lda, m.n_pad, m.Winv.get(), false, m.Wpanel.get(), extra);
    DevBuf<double> dmean(c, mt), colsq(c, mt), dvar(c, mt);
    {
        Timed t(c, c.main, "mean");
        k::rowdot(c.main, rows, m.lda, m.n_pad, mt, yrow, dmean.get(), false);
    }
    {
        Timed t(c, c.main, "variance_rows");
        k::row_sumsq(c.main, rows, m.lda, mt, m.n_pad, colsq.get());
        k::variance_finish(c.main, colsq.get(), m.sek.vertical, mt, dvar.get());
    }
    GP_HIP(hipMemcpyAsync(mean, dmean.get(), sizeof(double) * mt, hipMemcpyDefault, c.main));
    GP_HIP(hipMemcpyAsync(var, dvar.get(), sizeof(double) * mt, hipMemcpyDefault, c.main));
    c.sync();
}

FactorView view(const Model &m)
{
    FactorView v{ m.ctx, m.x.get(), m.n, m.n_pad, m.r, m.lda, m.sek, m.L.get(), m.Winv.get(), m.alpha.get(),
                  m.precision == GPRAT_PRECISION_MIXED, m.Wpanel.get(), m.wblock };
    v.refine_variance = m.refine_variance;
    if (m.precision == GPRAT_PRECISION_MIXED)
    {
        v.Kup = m.L.get();  // mixed models keep the covariance in the factor's strict upper triangle
        v.ldk = m.lda;
    }
    return v;
}

void predict(Model &m, const double *z, std::size_t z_len, int mt, double *mean, double *var)
{
    FactorView v = view(m);
    v.variance_correction = &m.variance_correction;
    predict_view(v, z, z_len, mt, mean, var);
}

// ---------------------------------------------------------------------------
// Batched fit of many GPs

static void check_desc(const gprat_gp_desc &d, int g)
{
    require(d.n >= 1 && d.n_reg >= 1, "batch: GP " + std::to_string(g) + ": n and n_reg must be >= 1");
    require(d.train_in != nullptr && d.train_in_len >= static_cast<std::size_t>(d.n) + d.n_reg - 1,
            "batch: GP " + std::to_string(g) + ": training input must hold n + n_reg - 1 values");
    require(d.train_out != nullptr && d.train_out_len >= static_cast<std::size_t>(d.n),
            "batch: GP " + std::to_string(g) + ": training output must hold n values");
}

void batch_fit(Context &c, Batch &b, int count, const gprat_gp_desc *gps, int precision)
{
    require(count >= 1 && gps != nullptr, "batch: need at least one GP");
    require(precision == GPRAT_PRECISION_FP64 || precision == GPRAT_PRECISION_MIXED, "batch: unknown precision mode");
    const bool mixed = precision == GPRAT_PRECISION_MIXED;
    c.make_current();
    b.ctx = &c;
    b.count = count;
    b.precision = precision;
    b.gps.resize(count);
    // arena layout: [inputs of all GPs][factors, inverses, alpha (and the
    // refinement work vector) of all GPs]
    std::vector<long> x_off(count), y_off(count), l_off(count), w_off(count), a_off(count), r_off(count, -1);
    long in_total = 0, out_total = 0;
    for (int g = 0; g < count; g++)
    {
        check_desc(gps[g], g);
        const long np = round_up(gps[g].n, 64);
        x_off[g] = in_total;
        in_total += round_up(gps[g].n + gps[g].n_reg - 1, 2);
        y_off[g] = in_total;
        in_total += np;
    }
    for (int g = 0; g < count; g++)
    {
        const long np = round_up(gps[g].n, 64);
        l_off[g] = in_total + out_total;
        out_total += np * np;
        w_off[g] = in_total + out_total;
        out_total += np * 64;
        a_off[g] = in_total + out_total;
        out_total += np;
        if (mixed)
        {
            r_off[g] = in_total + out_total;
            out_total += np;
        }
    }
    std::vector<double> staging(static_cast<std::size_t>(in_total), 0.0);
    for (int g = 0; g < count; g++)
    {
        std::copy(gps[g].train_in, gps[g].train_in + gps[g].n + gps[g].n_reg - 1, staging.begin() + x_off[g]);
        std::copy(gps[g].train_out, gps[g].train_out + gps[g].n, staging.begin() + y_off[g]);
    }
    b.arena = DevBuf<double>(c, static_cast<std::size_t>(in_total + out_total));
    double *base = b.arena.get();
    for (int g = 0; g < count; g++)
    {
        const gprat_gp_desc &d = gps[g];
        k::BatchGP &q = b.gps[g];
        q = k::BatchGP{};
        q.x = base + x_off[g];
        q.y = base + y_off[g];
        q.L = base + l_off[g];
        q.W = base + w_off[g];
        q.alpha = base + a_off[g];
        q.work = mixed ? base + r_off[g] : nullptr;
        q.n = d.n;
        q.n_pad = static_cast<int>(round_up(d.n, 64));
        q.n_reg = d.n_reg;
        q.lengthscale = d.theta[0];
        q.vertical = d.theta[1];
        q.noise = d.theta[2];
    }
    GP_HIP(hipMemcpyAsync(base, staging.data(), staging.size() * sizeof(double), hipMemcpyDefault, c.main));
    b.losses.assign(count, 0.0);
    b.refine_steps.assign(count, 0);
    const std::vector<int> info = batch_run(c, b.gps, b.losses.data(), mixed, b.refine_steps.data(), &b.gp_mixed);
    for (int g = 0; g < count; g++)
        if (info[g] != INT_MAX)
            throw Error(GPRAT_ERR_NOT_SPD, "batch: GP " + std::to_string(g)
                                               + ": covariance matrix is not positive definite (first non-positive pivot at column "
                                               + std::to_string(info[g]) + ")");
}

void batch_losses(Batch &b, double *losses) { std::copy(b.losses.begin(), b.losses.end(), losses); }

void batch_predict(Batch &b, int index, const double *z, std::size_t z_len, int mt, double *mean, double *var)
{
    require(index >= 0 && index < b.count, "batch: GP index out of range");
    const k::BatchGP &q = b.gps[index];
    const bool mixed = b.precision == GPRAT_PRECISION_MIXED && (b.gp_mixed.empty() || b.gp_mixed[index]);
    predict_view(FactorView{ b.ctx, q.x, q.n, q.n_pad, q.n_reg, q.n_pad, SEK{ q.lengthscale, q.vertical, q.noise }, q.L, q.W, q.alpha,
                             mixed },
                 z, z_len, mt, mean, var);
}

double loss(Model &m)
{
    Context &c = *m.ctx;
    c.make_current();
    k::loss_terms(c.main, m.L.get(), m.lda, m.y.get(), m.alpha.get(), m.n, c.scalars.get());
    double t[3];
    GP_HIP(hipMemcpyAsync(t, c.scalars.get(), sizeof t, hipMemcpyDefault, c.main));
    c.sync();
    // core/src/cpu/gp_optimizer.cpp:205-218
    const double Nn = static_cast<double>(m.n);
    double l = t[0] + t[1];
    l += Nn * std::log(2.0 * std::numbers::pi);
    return 0.5 * l / Nn;
}

// Softplus chain factors, as core/src/cpu/gp_optimizer.cpp:12-36.
static double to_unconstrained(double p, bool noise)
{
    return noise ? std::log(std::exp(p - 1e-6) - 1.0) : std::log(std::exp(p) - 1.0);
}
static double sigmoid(double p) { return 1.0 / (1.0 + std::exp(-p)); }

// Loss and softplus-space gradient at theta. K^-1 is formed as W^T W with
// W = L^-1 (recursive TRSM on the triangular identity, then one triangular
// MFMA product), n^3 flops in total instead of the reference's explicit
// 2 n^3 double solve (core/src/cpu/gp_functions.cpp:953-966); the trace and
// quadratic terms of core/src/cpu/tiled_algorithms.cpp:296-426 then come from
// one fused pass over the lower triangle of K^-1.
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
                   double grad[3])
{
    Model m;
    fit(c, m, x, x_len, y, y_len, n, r, theta, GPRAT_PRECISION_FP64);
    model_gradient(m, trainable, loss_out, grad);
}

void model_gradient(Model &m, const int trainable[3], double *loss_out, double grad[3])
{
    Context &c = *m.ctx;
    c.make_current();
    const int n = m.n, r = m.r;
    k::loss_terms(c.main, m.L.get(), m.lda, m.y.get(), m.alpha.get(), m.n, c.scalars.get());
    double t[3];
    GP_HIP(hipMemcpyAsync(t, c.scalars.get(), sizeof t, hipMemcpyDefault, c.main));
    c.sync();
    const double Nn = static_cast<double>(n);
    {
        double l = t[0] + t[1];
        l += Nn * std::log(2.0 * std::numbers::pi);
        *loss_out = 0.5 * l / Nn;
    }
    grad[0] = grad[1] = grad[2] = 0.0;
    if (!(trainable[0] || trainable[1] || trainable[2]))
        return;
    const std::size_t elems = static_cast<std::size_t>(m.n_pad) * m.lda;
    DevBuf<double> W(c, elems);
    DevBuf<double> Kinv(c, elems);
    k::fill(c.main, W.get(), static_cast<long>(elems), 0.0);
    {
        // the identity right-hand side lives in Kinv until the product overwrites it
        k::fill(c.main, Kinv.get(), static_cast<long>(elems), 0.0);
        k::set_diagonal(c.main, Kinv.get(), m.lda, m.n_pad, 1.0);
    }
    {
        Timed tm(c, c.main, "inverse");
        trsm_blocked(c, m.L.get(), m.lda, m.Winv.get(), m.Wpanel.get(), m.wblock, m.n_pad, Kinv.get(), m.lda, m.n_pad, true, W.get(),
                     m.lda);
        k::GemmArgs g{};
        g.A = W.get();
        g.lda = m.lda;
        g.B = W.get();
        g.ldb = m.lda;
        g.C = Kinv.get();
        g.ldc = m.lda;
        g.M = g.N = g.K = m.n_pad;
        g.a_trans = true;
        g.b_trans = false;
        g.lower = true;
        g.k_from_tile = true;
        g.set = true;
        g.role = k::GemmRole::Inverse;
        k::gemm_f64(c.main, g);
    }
    W.reset();
    const double sig_l = sigmoid(to_unconstrained(m.sek.lengthscale, false));
    const double sig_v = sigmoid(to_unconstrained(m.sek.vertical, false));
    const double sig_n = sigmoid(to_unconstrained(m.sek.noise, true));
    DevBuf<double> part(c, static_cast<std::size_t>(k::grad_tiles(n)) * 5);
    {
        Timed tm(c, c.main, "gradient");
        k::gradient_sums(c.main, Kinv.get(), m.lda, m.x.get(), m.alpha.get(), r, m.sek, sig_l, sig_v, n, part.get(),
                         c.scalars.get() + 8);
    }
    double sums[5];
    GP_HIP(hipMemcpyAsync(sums, c.scalars.get() + 8, sizeof sums, hipMemcpyDefault, c.main));
    c.sync();
    // core/src/cpu/gp_optimizer.cpp:222-225 and tiled_algorithms.cpp:413-426
    if (trainable[0])
        grad[0] = 0.5 / Nn * (sums[0] - sums[2]);
    if (trainable[1])
        grad[1] = 0.5 / Nn * (sums[1] - sums[3]);
    if (trainable[2])
        grad[2] = sig_n * (0.5 / Nn * (sums[4] - t[2]));
}

void cholesky_tiles(Model &m, int T, int b, double *out)
{
    Context &c = *m.ctx;
    c.make_current();
    require(static_cast<long>(T) * b == m.n, "cholesky_tiles: tiling does not match the model");
    const std::size_t count = static_cast<std::size_t>(T) * (T + 1) / 2 * b * b;
    DevBuf<double> d(c, count);
    k::gather_tiles(c.main, m.L.get(), m.lda, m.x.get(), m.r, m.sek, T, b, d.get());
    GP_HIP(hipMemcpyAsync(out, d.get(), count * sizeof(double), hipMemcpyDefault, c.main));
    c.sync();
}

}  // namespace gpamd
