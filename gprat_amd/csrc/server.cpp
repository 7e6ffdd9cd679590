// Request loop with continuous batching (see server.hpp for the policy).
#include "server.hpp"

#include <algorithm>
#include <climits>

namespace gpamd
{

namespace
{
double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

int status_of(const std::exception &e)
{
    if (auto *g = dynamic_cast<const Error *>(&e))
        return g->status;
    if (dynamic_cast<const std::bad_alloc *>(&e))
        return GPRAT_ERR_OUT_OF_MEMORY;
    return GPRAT_ERR_INTERNAL;
}
}  // namespace

Server::Server(const ServerConfig &cfg) : cfg_(cfg)
{
    require(cfg_.slots >= 1 && cfg_.slots <= 16, "server: slots must be in [1, 16]");
    require(cfg_.precision == GPRAT_PRECISION_FP64 || cfg_.precision == GPRAT_PRECISION_MIXED, "server: unknown precision mode");
    require(cfg_.batch_max_n >= 0 && cfg_.batch_max >= 1, "server: bad batching limits");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    {
        (void)hipGetLastError();
        throw Error(GPRAT_ERR_NO_DEVICE, "no HIP device available");
    }
    require(cfg_.device >= 0 && cfg_.device < count, "server: device index out of range");
    GP_HIP(hipSetDevice(cfg_.device));
    std::size_t free_b = 0, total_b = 0;
    GP_HIP(hipMemGetInfo(&free_b, &total_b));
    if (cfg_.budget == 0)
        cfg_.budget = total_b / 10 * 7;
    // Blocks cached by the slots' pools live outside the budget and are kept
    // (freeing synchronises the whole device). A slot whose own cache has no
    // block of the right size takes one from another slot's cache (each
    // slot's context runs on one stream); only when the device is full is
    // every cache freed and the allocation tried again.
    st_.budget = cfg_.budget;
    for (int s = 0; s < cfg_.slots; s++)
        ctx_.push_back(std::make_unique<Context>(cfg_.device, 1));
    std::vector<DevicePool *> pools;
    for (auto &c : ctx_)
        pools.push_back(&c->pool);
    for (auto &c : ctx_)
    {
        c->pool.share_with(pools, c->main);
        c->pool.on_out_of_memory = [this]
        {
            cache_flushes_++;
            for (auto &o : ctx_)
                o->pool.trim(0);
        };
    }
    for (int s = 0; s < cfg_.slots; s++)
        workers_.emplace_back([this, s] { worker(s); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    work_cv_.notify_all();
    for (auto &t : workers_)
        t.join();
    models_.clear();
    ctx_.clear();
}

// ---------------------------------------------------------------------------
// Submission

long long Server::submit_fit(const gprat_gp_desc &d)
{
    require(d.n >= 1 && d.n_reg >= 1, "fit: n and n_reg must be >= 1");
    require(d.train_in != nullptr && d.train_in_len >= static_cast<std::size_t>(d.n) + d.n_reg - 1,
            "fit: training input must hold n + n_reg - 1 values");
    require(d.train_out != nullptr && d.train_out_len >= static_cast<std::size_t>(d.n), "fit: training output must hold n values");
    auto e = std::make_unique<Entry>();
    e->n = d.n;
    e->r = d.n_reg;
    std::copy(d.theta, d.theta + 3, e->theta);
    e->small = cfg_.batch_max_n > 0 && d.n <= cfg_.batch_max_n;
    e->x.assign(d.train_in, d.train_in + d.n + d.n_reg - 1);
    e->y.assign(d.train_out, d.train_out + d.n);
    Request r;
    r.kind = 0;
    r.submitted = Clock::now();
    long long ticket;
    {
        std::lock_guard<std::mutex> g(mu_);
        ticket = next_ticket_++;
        r.ticket = ticket;
        r.model = ticket;
        e->id = ticket;
        models_.emplace(ticket, std::move(e));
        live_.insert(ticket);
        queue_.push_back(std::move(r));
        st_.submitted++;
    }
    work_cv_.notify_one();
    return ticket;
}

long long Server::submit_predict(long long model, const double *z, std::size_t z_len, int m, double *mean, double *var)
{
    require(m >= 1, "predict: need at least one test point");
    require(mean != nullptr, "predict: mean output is null");
    Request r;
    r.kind = 1;
    r.model = model;
    r.m = m;
    r.mean = mean;
    r.var = var;
    r.submitted = Clock::now();
    long long ticket;
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = models_.find(model);
        require(it != models_.end() && !it->second->released, "predict: unknown or released model " + std::to_string(model));
        const int reg = it->second->r;
        require(z != nullptr && z_len >= static_cast<std::size_t>(m) + reg - 1, "predict: test input must hold m + n_reg - 1 values");
        r.z.assign(z, z + m + reg - 1);
        ticket = next_ticket_++;
        r.ticket = ticket;
        live_.insert(ticket);
        queue_.push_back(std::move(r));
        st_.submitted++;
    }
    work_cv_.notify_one();
    return ticket;
}

bool Server::wait(long long ticket, double timeout_ms, RequestResult &out)
{
    std::unique_lock<std::mutex> lk(mu_);
    require(live_.count(ticket) > 0, "wait: unknown or already retired ticket " + std::to_string(ticket));
    auto ready = [&] { return done_.count(ticket) > 0; };
    if (timeout_ms < 0)
        done_cv_.wait(lk, ready);
    else if (!done_cv_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
        return false;
    auto it = done_.find(ticket);
    out = std::move(it->second);
    done_.erase(it);
    live_.erase(ticket);
    return true;
}

void Server::release_model(long long model)
{
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = models_.find(model);
        require(it != models_.end() && !it->second->released, "release: unknown or released model " + std::to_string(model));
        Entry &e = *it->second;
        e.released = true;
        if (e.pins == 0 && !e.rebuilding && e.state != State::Pending)
        {
            resident_ -= e.bytes;
            models_.erase(it);
        }
    }
    work_cv_.notify_all();
}

ServerStats Server::stats()
{
    std::lock_guard<std::mutex> g(mu_);
    ServerStats s = st_;
    s.cache_flushes = cache_flushes_.load();
    for (const auto &c : ctx_)
        s.borrowed_blocks += c->pool.borrowed();
    s.resident_bytes = resident_;
    s.reserved_bytes = reserved_;
    s.resident_models = 0;
    for (const auto &kv : models_)
        if (kv.second->model)
            s.resident_models++;
    return s;
}

// ---------------------------------------------------------------------------
// Memory estimates (bytes a request may allocate while it runs, including
// the factor it leaves resident)

std::size_t Server::fit_need(const Entry &e) const
{
    const std::size_t n = e.n, r = e.r;
    std::size_t np, ld, temp;
    if (e.batched_path)
    {
        np = round_up(e.n, 64);
        ld = np;
        temp = 4096;
    }
    else
    {
        np = padded(e.n);
        ld = leading_dim(static_cast<long>(np));
        // Cholesky panel copy and panel scratch (panel width x n_pad doubles
        // each), the diagonal-block scratch, solve and refinement vectors,
        // the refinement's partial sums
        const std::size_t bo = static_cast<std::size_t>(panel_width(static_cast<int>(np)));
        temp = 8 * (3 * bo * np + 8 * np + static_cast<std::size_t>(k::mean_splits(e.n, e.n)) * n);
    }
    // factor, 64-block inverses, panel inverses (panel width wide), vectors
    const std::size_t wide = e.batched_path ? 0 : np * static_cast<std::size_t>(panel_width(static_cast<int>(np)));
    const std::size_t resident = 8 * (np * ld + np * 64 + wide + 2 * np + n + r);
    return resident + temp + (std::size_t(1) << 20);
}

std::size_t Server::predict_need(const Entry &e, int m) const
{
    const std::size_t np = e.batched_path ? round_up(e.n, 64) : padded(e.n);
    const std::size_t mp = round_up(m, 64);
    const std::size_t blocks = (np + 255) / 256;
    const std::size_t splits = static_cast<std::size_t>(k::mean_splits(e.n, m));
    // K*^T, the staged V block and the panel copy (panel width wide), block
    // inverses built for a factor without kept ones, the mean's partial sums
    const std::size_t bo = e.batched_path ? 256 : static_cast<std::size_t>(panel_width(static_cast<int>(np)));
    std::size_t bytes = 8 * (np * mp + bo * mp + 4 * mp + blocks * 65536 + bo * np + splits * m + m + e.r) + (std::size_t(1) << 20);
    if (!e.small && cfg_.precision == GPRAT_PRECISION_MIXED)
    {
        // FP64-refined variance: FP32 K*^T / solution, FP64 K*^T, solution
        // and residual product, FP32 block inverses and their transposes,
        // the FP32 panel and row-block copies and the staged blocks
        bytes += 28 * np * mp + 24 * np * bo + 8 * bo * mp;
    }
    return bytes;
}

// Estimated flops of a request: the factorisation n^3/3 for a fit (or a
// predict that rebuilds an evicted factor) plus n^2 m for a predict's
// variance solve.
double Server::cost(const Request &r, const Entry &e) const
{
    const double n = e.n;
    double f = (r.kind == 0 || !e.model) ? n * n * n / 3.0 : 0.0;
    if (r.kind == 1)
        f += n * n * r.m;
    return f;
}

// Cost model of the small-fit routing (milliseconds, FP64, n <= 2048),
// measured on MI355X (host-timed, n = 256 .. 2048):
// * a single-GP fit uses the whole device but is bound by its launches'
//   latency, about linear in n (0.27 ms at n = 256, 1.89 ms at 2048);
// * a batched fit runs its GPs above n_pad 128 as level-synchronous launches
//   (two per 64-row block of the largest GP), so it lasts as long as that
//   chain of launches (0.34 ms for one GP of n = 256, 1.55 ms at 1024,
//   4.67 ms at 2048) plus the GPs' work at the batch's throughput (16 GPs of
//   n = 2048: 5.51 ms).
static double single_fit_ms(int n) { return 0.05 + 0.00092 * n; }
static double batch_chain_ms(int n_max)
{
    const double n = n_max;
    return 0.14 + 5.37e-4 * n + 8.0e-7 * n * n;
}
static double batch_work_ms(int n)
{
    const double f = static_cast<double>(n) * n * n / 3.0;
    return f / 4.0e10;  // 40 TF/s
}

// Largest n of the small fits that should share one batched launch now,
// from the small fits waiting at and after `first` (up to batch_max), or 0
// for none. For each cut n_c, the fits with n <= n_c would take one batched
// fit (batch_chain_ms(n_c) plus their work) against their single-GP times
// spread over the slots; the cut with the largest saving wins (at least two
// fits). A lone small fit, or a few small fits that the slots finish
// sooner one by one, take the single-GP path.
int Server::batch_cut(std::deque<Request>::iterator first)
{
    std::vector<int> ns;
    for (auto it = first; it != queue_.end() && static_cast<int>(ns.size()) < cfg_.batch_max; ++it)
    {
        if (it->kind != 0)
            continue;
        auto qm = models_.find(it->model);
        if (qm != models_.end() && qm->second->small)
            ns.push_back(qm->second->n);
    }
    std::sort(ns.begin(), ns.end());
    double single = 0.0, work = 0.0, best = 0.0;
    int cut = 0;
    for (std::size_t k = 0; k < ns.size(); k++)
    {
        single += single_fit_ms(ns[k]);
        work += batch_work_ms(ns[k]);
        const int count = static_cast<int>(k) + 1;
        if (count < 2 || (k + 1 < ns.size() && ns[k + 1] == ns[k]))
            continue;
        const double saving = single / std::min(count, cfg_.slots) - (batch_chain_ms(ns[k]) + work);
        if (saving > best)
        {
            best = saving;
            cut = ns[k];
        }
    }
    return cut;
}

// Fits the budget, evicting idle resident factors least recently used first.
bool Server::admit(std::size_t need)
{
    while (resident_ + reserved_ + need > cfg_.budget)
    {
        Entry *victim = nullptr;
        for (auto &kv : models_)
        {
            Entry &e = *kv.second;
            if (e.model && e.pins == 0 && !e.rebuilding && (!victim || e.last_use < victim->last_use))
                victim = &e;
        }
        if (!victim)
            return false;
        resident_ -= victim->bytes;
        victim->bytes = 0;
        victim->model.reset();
        st_.evictions++;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Scheduling

void Server::complete(const Request &r, RequestResult res)
{
    res.kind = r.kind;
    res.model = r.model;
    if (res.status != GPRAT_OK)
        st_.failed++;
    st_.completed++;
    if (r.kind == 0)
        st_.fits++;
    else
        st_.predicts++;
    done_[r.ticket] = std::move(res);
}

// Takes the next admissible work in arrival order. Predicts whose model is
// still being fitted (or rebuilt) wait without blocking later requests; a
// request that does not fit the budget blocks the ones behind it until
// memory frees, so large requests are not starved.
bool Server::pick(Job &job)
{
    for (auto it = queue_.begin(); it != queue_.end();)
    {
        Request &r = *it;
        auto mit = models_.find(r.model);
        if (mit == models_.end() || (r.kind == 1 && mit->second->released))
        {
            RequestResult res;
            res.status = GPRAT_ERR_ARGUMENT;
            res.error = "model " + std::to_string(r.model) + " was released";
            complete(r, std::move(res));
            it = queue_.erase(it);
            done_cv_.notify_all();
            continue;
        }
        Entry &e = *mit->second;
        std::size_t need;
        int cut = 0;
        if (r.kind == 1)
        {
            if (e.state == State::Pending || e.rebuilding)
            {
                ++it;
                continue;
            }
            if (e.state == State::Failed)
            {
                RequestResult res;
                res.status = e.status;
                res.error = "model " + std::to_string(r.model) + " failed to fit: " + e.error;
                complete(r, std::move(res));
                it = queue_.erase(it);
                done_cv_.notify_all();
                continue;
            }
            need = predict_need(e, r.m) + (e.model ? 0 : fit_need(e));
        }
        else
        {
            // a small fit's path is chosen now, from what is waiting with it
            if (e.small)
            {
                cut = batch_cut(it);
                e.batched_path = e.n <= cut;
            }
            need = fit_need(e);
        }
        const bool heavy = cost(r, e) > cfg_.heavy_flops && !(r.kind == 0 && e.batched_path);
        if (heavy && heavy_in_flight_ >= std::max(1, cfg_.slots - 1))
        {
            ++it;  // keeps its place; short requests behind it may start
            continue;
        }
        // a predict's own factor must not be the one its admission evicts
        // (need assumed it resident): it is pinned first
        if (r.kind == 1)
            e.pins++;
        const bool fits = admit(need);
        if (r.kind == 1 && !fits)
            e.pins--;
        if (!fits)
        {
            if (in_flight_ > 0)
                return false;
            RequestResult res;
            res.status = GPRAT_ERR_OUT_OF_MEMORY;
            res.error = "request needs " + std::to_string(need) + " bytes, more than the HBM budget of "
                        + std::to_string(cfg_.budget) + " bytes allows";
            if (r.kind == 0)
            {
                e.state = State::Failed;
                e.status = res.status;
                e.error = res.error;
            }
            complete(r, std::move(res));
            it = queue_.erase(it);
            done_cv_.notify_all();
            continue;
        }
        const auto now = Clock::now();
        r.need = need;
        r.admitted = now;
        r.heavy = heavy;
        heavy_in_flight_ += heavy ? 1 : 0;
        reserved_ += need;
        if (r.kind == 1)
        {
            r.rebuild = !e.model;
            if (r.rebuild)
                e.rebuilding = true;
        }
        const bool gather = r.kind == 0 && e.batched_path;
        job.reqs.push_back(std::move(r));
        it = queue_.erase(it);
        // continuous batching: take every other small fit already waiting
        while (gather && it != queue_.end() && static_cast<int>(job.reqs.size()) < cfg_.batch_max)
        {
            Request &q = *it;
            auto qm = models_.find(q.model);
            if (q.kind != 0 || qm == models_.end() || !qm->second->small || qm->second->n > cut)
            {
                ++it;
                continue;
            }
            qm->second->batched_path = true;
            const std::size_t qn = fit_need(*qm->second);
            if (!admit(qn))
            {
                qm->second->batched_path = false;
                break;
            }
            q.need = qn;
            q.admitted = now;
            reserved_ += qn;
            job.reqs.push_back(std::move(q));
            it = queue_.erase(it);
        }
        st_.peak_bytes = std::max(st_.peak_bytes, resident_ + reserved_);
        return true;
    }
    return false;
}

void Server::worker(int slot)
{
    Context &c = *ctx_[slot];
    c.make_current();
    for (;;)
    {
        Job job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;)
            {
                if (pick(job))
                    break;
                if (stop_ && in_flight_ == 0)
                {
                    // nothing can make progress any more: fail what is left
                    while (!queue_.empty())
                    {
                        RequestResult res;
                        res.status = GPRAT_ERR_ARGUMENT;
                        res.error = "server shut down before the request ran";
                        complete(queue_.front(), std::move(res));
                        queue_.pop_front();
                    }
                    done_cv_.notify_all();
                    return;
                }
                work_cv_.wait(lk);
            }
            in_flight_++;
        }
        std::vector<RequestResult> res(job.reqs.size());
        run(c, job, res);
        {
            std::lock_guard<std::mutex> g(mu_);
            finish(job, res);
            in_flight_--;
        }
        done_cv_.notify_all();
        work_cv_.notify_all();
    }
}

// ---------------------------------------------------------------------------
// Execution (no lock held; entries are pinned or pending, so nothing else
// modifies them)

void Server::build_single(Context &c, Job &job, Entry &e)
{
    auto m = std::make_unique<Model>();
    // small fits are FP64 on either path (mixed precision gains nothing at
    // sizes bound by launch latency, and both paths then agree to rounding)
    fit(c, *m, e.x.data(), e.x.size(), e.y.data(), e.y.size(), e.n, e.r, e.theta, e.small ? GPRAT_PRECISION_FP64 : cfg_.precision);
    job.built.emplace(e.id, std::move(m));
}

void Server::build_batched(Context &c, Job &job, std::vector<Entry *> &es, std::vector<RequestResult *> &res)
{
    std::vector<std::unique_ptr<Model>> ms;
    std::vector<k::BatchGP> gps;
    for (Entry *e : es)
    {
        auto m = std::make_unique<Model>();
        m->ctx = &c;
        m->n = e->n;
        m->r = e->r;
        m->n_pad = static_cast<int>(round_up(e->n, 64));
        m->lda = m->n_pad;
        m->sek = SEK{ e->theta[0], e->theta[1], e->theta[2] };
        m->precision = GPRAT_PRECISION_FP64;
        const std::size_t np = m->n_pad;
        m->x = DevBuf<double>(c, e->x.size());
        m->y = DevBuf<double>(c, np);
        m->L = DevBuf<double>(c, np * np);
        m->Winv = DevBuf<double>(c, np * 64);
        m->alpha = DevBuf<double>(c, np);
        GP_HIP(hipMemcpyAsync(m->x.get(), e->x.data(), e->x.size() * sizeof(double), hipMemcpyDefault, c.main));
        k::fill(c.main, m->y.get(), static_cast<long>(np), 0.0);
        GP_HIP(hipMemcpyAsync(m->y.get(), e->y.data(), e->y.size() * sizeof(double), hipMemcpyDefault, c.main));
        k::BatchGP q{};
        q.x = m->x.get();
        q.y = m->y.get();
        q.L = m->L.get();
        q.W = m->Winv.get();
        q.alpha = m->alpha.get();
        q.n = m->n;
        q.n_pad = m->n_pad;
        q.n_reg = m->r;
        q.lengthscale = e->theta[0];
        q.vertical = e->theta[1];
        q.noise = e->theta[2];
        gps.push_back(q);
        ms.push_back(std::move(m));
    }
    std::vector<double> losses(es.size());
    const std::vector<int> info = batch_run(c, gps, losses.data());
    for (std::size_t i = 0; i < es.size(); i++)
    {
        if (res[i])
        {
            res[i]->batched = static_cast<int>(es.size());
            res[i]->loss = losses[i];
        }
        if (info[i] != INT_MAX)
        {
            if (res[i])
            {
                res[i]->status = GPRAT_ERR_NOT_SPD;
                res[i]->error = "covariance matrix is not positive definite (first non-positive pivot at column "
                                + std::to_string(info[i]) + ")";
            }
            continue;
        }
        job.built.emplace(es[i]->id, std::move(ms[i]));
    }
}

void Server::run(Context &c, Job &job, std::vector<RequestResult> &res)
{
    // entries of this job: pinned or pending, so the map nodes stay put
    std::vector<Entry *> es;
    {
        std::lock_guard<std::mutex> g(mu_);
        for (const auto &r : job.reqs)
            es.push_back(models_.at(r.model).get());
    }
    const Request &r0 = job.reqs[0];
    try
    {
        if (r0.kind == 0)
        {
            if (es[0]->batched_path)
            {
                std::vector<RequestResult *> rp;
                for (auto &x : res)
                    rp.push_back(&x);
                build_batched(c, job, es, rp);
            }
            else
            {
                build_single(c, job, *es[0]);
                res[0].loss = loss(*job.built.at(es[0]->id));
            }
        }
        else
        {
            Entry &e = *es[0];
            const Model *m = e.model.get();
            if (r0.rebuild)
            {
                if (e.batched_path)
                {
                    std::vector<Entry *> one{ &e };
                    std::vector<RequestResult *> none{ nullptr };
                    build_batched(c, job, one, none);
                }
                else
                {
                    build_single(c, job, e);
                }
                auto b = job.built.find(e.id);
                require(b != job.built.end(), "predict: rebuilding the evicted factor failed (not positive definite)");
                m = b->second.get();
                res[0].refit = 1;
            }
            FactorView v = view(*m);
            v.ctx = &c;  // this slot's stream and pool
            predict_view(v, r0.z.data(), r0.z.size(), r0.m, r0.mean, r0.var);
        }
    }
    catch (const std::exception &ex)
    {
        for (auto &x : res)
            if (x.status == GPRAT_OK)
            {
                x.status = status_of(ex);
                x.error = ex.what();
            }
        try
        {
            c.sync();
        }
        catch (...)
        {
        }
    }
    const auto now = Clock::now();
    for (std::size_t i = 0; i < job.reqs.size(); i++)
    {
        res[i].queue_ms = ms_between(job.reqs[i].submitted, job.reqs[i].admitted);
        res[i].service_ms = ms_between(job.reqs[i].admitted, now);
    }
}

void Server::finish(Job &job, std::vector<RequestResult> &res)
{
    if (job.reqs[0].kind == 0 && res[0].batched > 0)
        st_.batch_launches++;
    for (std::size_t i = 0; i < job.reqs.size(); i++)
    {
        Request &r = job.reqs[i];
        reserved_ -= r.need;
        heavy_in_flight_ -= r.heavy ? 1 : 0;
        auto mit = models_.find(r.model);
        Entry &e = *mit->second;
        auto b = job.built.find(e.id);
        if (b != job.built.end())
        {
            // a fitted or rebuilt factor becomes resident
            e.model = std::move(b->second);
            e.bytes = e.model->bytes();
            resident_ += e.bytes;
            job.built.erase(b);
            if (r.kind == 1)
                st_.refits++;
        }
        if (r.kind == 0)
        {
            e.state = res[i].status == GPRAT_OK ? State::Ready : State::Failed;
            e.status = res[i].status;
            e.error = res[i].error;
            if (res[i].batched)
                st_.batched_fits++;
        }
        else
        {
            e.pins--;
            if (r.rebuild)
                e.rebuilding = false;
        }
        e.last_use = ++clock_;
        if (e.released && e.pins == 0 && !e.rebuilding)
        {
            resident_ -= e.bytes;
            models_.erase(mit);
        }
        complete(r, std::move(res[i]));
    }
    job.built.clear();
}

}  // namespace gpamd
