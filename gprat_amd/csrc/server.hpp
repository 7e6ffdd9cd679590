// Request loop with continuous batching over one GPU.
//
// Fit and predict requests are queued by the caller and picked up by a fixed
// number of slots; each slot is a worker thread with its own Context (one
// stream, its own caching pool), so requests admitted together overlap on
// the device. A request is admitted when the HBM budget still holds its
// working memory next to every resident factor and every other admitted
// request; otherwise idle resident factors are evicted least recently used
// first, and a predict against an evicted model rebuilds its factor from
// the host copy of its data, the same way it was first built. Small fits
// waiting in the queue are taken together into one persistent batched
// launch. All kernels reduce in a fixed order, so a request's result does
// not depend on what else is in flight.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "engine.hpp"

namespace gpamd
{

struct ServerConfig
{
    int device = 0;
    int slots = 4;
    std::size_t budget = 0;  // bytes; 0 = 70% of the device's memory
    int precision = GPRAT_PRECISION_MIXED;
    int batch_max_n = 2048;  // fits with n <= batch_max_n run in batched launches
    int batch_max = 256;     // fits per batched launch at most
    // Requests estimated above heavy_flops run on at most max(1, slots - 1)
    // slots at once, so one slot always stays free for short requests.
    double heavy_flops = 2e11;
};

struct RequestResult
{
    int status = GPRAT_OK;
    int kind = 0;  // 0 fit, 1 predict
    long long model = 0;
    double loss = 0.0;
    double queue_ms = 0.0;    // submit -> admitted
    double service_ms = 0.0;  // admitted -> finished
    int refit = 0;            // predict: the factor was rebuilt after eviction
    int batched = 0;          // fit: GPs in the launch that fitted it (0: single path)
    std::string error;
};

struct ServerStats
{
    long long submitted = 0, completed = 0, failed = 0;
    long long fits = 0, predicts = 0, batched_fits = 0, batch_launches = 0;
    long long evictions = 0, refits = 0;
    long long resident_models = 0;
    std::size_t resident_bytes = 0, reserved_bytes = 0, peak_bytes = 0, budget = 0;
    long long cache_flushes = 0;    // device full: every slot's cached blocks freed
    long long borrowed_blocks = 0;  // cached blocks one slot took from another's cache
};

class Server
{
  public:
    explicit Server(const ServerConfig &cfg);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // The returned ticket of a fit is also the id of its model.
    long long submit_fit(const gprat_gp_desc &d);
    // mean/var (var may be null) are written when the request finishes and
    // must stay valid until wait() has returned it.
    long long submit_predict(long long model, const double *z, std::size_t z_len, int m, double *mean, double *var);
    // true: finished, `out` filled and the ticket retired; false: timeout.
    bool wait(long long ticket, double timeout_ms, RequestResult &out);
    void release_model(long long model);
    ServerStats stats();

  private:
    using Clock = std::chrono::steady_clock;
    enum class State
    {
        Pending,
        Ready,
        Failed
    };
    struct Entry
    {
        long long id = 0;
        int n = 0, r = 0;
        double theta[3] = { 0, 0, 0 };
        bool small = false;         // n <= batch_max_n: FP64, batched or single-GP path by load
        bool batched_path = false;  // built (and rebuilt) by the batched kernel (chosen when the fit starts)
        std::vector<double> x, y;   // host copies for rebuilding after eviction
        std::unique_ptr<Model> model;
        std::size_t bytes = 0;
        State state = State::Pending;
        int status = GPRAT_OK;
        std::string error;
        int pins = 0;
        bool rebuilding = false;
        bool released = false;
        unsigned long long last_use = 0;
    };
    struct Request
    {
        long long ticket = 0;
        int kind = 0;
        long long model = 0;
        std::vector<double> z;
        int m = 0;
        double *mean = nullptr, *var = nullptr;
        Clock::time_point submitted, admitted;
        std::size_t need = 0;
        bool rebuild = false;
        bool heavy = false;
    };
    struct Job
    {
        std::vector<Request> reqs;  // one predict, one fit, or several small fits
        std::map<long long, std::unique_ptr<Model>> built;  // factors made by this job, by model id
    };

    void worker(int slot);
    bool pick(Job &job);
    bool admit(std::size_t need);
    void run(Context &c, Job &job, std::vector<RequestResult> &res);
    void build_single(Context &c, Job &job, Entry &e);
    void build_batched(Context &c, Job &job, std::vector<Entry *> &es, std::vector<RequestResult *> &res);
    void finish(Job &job, std::vector<RequestResult> &res);
    void complete(const Request &r, RequestResult res);
    double cost(const Request &r, const Entry &e) const;
    int batch_cut(std::deque<Request>::iterator first);
    std::size_t fit_need(const Entry &e) const;
    std::size_t predict_need(const Entry &e, int m) const;

    ServerConfig cfg_;
    std::vector<std::unique_ptr<Context>> ctx_;
    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<Request> queue_;
    std::map<long long, std::unique_ptr<Entry>> models_;
    std::unordered_map<long long, RequestResult> done_;
    std::unordered_set<long long> live_;  // tickets issued and not yet retired by wait()
    std::vector<std::thread> workers_;
    long long next_ticket_ = 1;
    unsigned long long clock_ = 0;
    std::size_t resident_ = 0, reserved_ = 0;
    int in_flight_ = 0;
    int heavy_in_flight_ = 0;
    bool stop_ = false;
    ServerStats st_;
    std::atomic<long long> cache_flushes_{ 0 };
};

}  // namespace gpamd
