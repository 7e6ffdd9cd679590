// Host C++ facade: reference-shaped classes over the gfx950 C ABI. The
// optimiser's scalar side (softplus transforms, Adam) runs here, as in the
// reference; every O(n^2) or O(n^3) step runs on the GPU.
#include "gprat.hpp"

#include <mutex>
#include <thread>

#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "gprat_amd.h"

namespace
{

void check(int status, const char *what)
{
    if (status == GPRAT_OK)
        return;
    std::string msg = std::string(what) + ": " + gprat_last_error();
    if (status == GPRAT_ERR_ARGUMENT)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Parameter transforms and Adam ("option 2"), as
// core/src/cpu/gp_optimizer.cpp:12-36,154-183.
double to_constrained(double p, bool noise)
{
    return noise ? std::log(1.0 + std::exp(p)) + 1e-6 : std::log(1.0 + std::exp(p));
}

double to_unconstrained(double p, bool noise)
{
    return noise ? std::log(std::exp(p - 1e-6) - 1.0) : std::log(std::exp(p) - 1.0);
}

double adam_step(double u, const gprat_hyper::AdamParams &a, double m, double v, std::size_t iter)
{
    const double b1t = std::pow(a.beta1, static_cast<double>(iter + 1));
    const double b2t = std::pow(a.beta2, static_cast<double>(iter + 1));
    const double nu = a.learning_rate * std::sqrt(1.0 - b2t) / (1.0 - b1t);
    return u - nu * m / (std::sqrt(v) + a.epsilon);
}

// Moments and step for every trainable parameter, each with its gradient at
// the same theta (core/src/cpu/tiled_algorithms.cpp:428-447, called for l, v,
// noise in that order by core/src/cpu/gp_functions.cpp:986-1024).
void adam_update(gprat_hyper::SEKParams &p,
                 const std::vector<bool> &trainable,
                 const gprat_hyper::AdamParams &a,
                 const double grad[3],
                 std::size_t iter)
{
    for (std::size_t k = 0; k < 3; k++)
    {
        if (!trainable[k])
            continue;
        const bool noise = (k == 2);
        p.m_T[k] = a.beta1 * p.m_T[k] + (1.0 - a.beta1) * grad[k];
        p.w_T[k] = a.beta2 * p.w_T[k] + (1.0 - a.beta2) * grad[k] * grad[k];
        const double u = to_unconstrained(p.get_param(k), noise);
        p.set_param(k, to_constrained(adam_step(u, a, p.m_T[k], p.w_T[k], iter), noise));
    }
}

int g_host_threads = 0;
enum class RuntimeState
{
    NotStarted,
    Running,
    Suspended,
    Stopped
};
RuntimeState g_runtime_state = RuntimeState::NotStarted;
std::mutex g_runtime_mu;

}  // namespace

namespace gprat_hyper
{

SEKParams::SEKParams(double l, double v, double n) : lengthscale(l), vertical_lengthscale(v), noise_variance(n)
{
    m_T.resize(size());
    w_T.resize(size());
}

std::size_t SEKParams::size() { return 3; }

void SEKParams::set_param(std::size_t index, double value)
{
    if (index == 0)
        lengthscale = value;
    else if (index == 1)
        vertical_lengthscale = value;
    else if (index == 2)
        noise_variance = value;
    else
        throw std::invalid_argument("Set Invalid param_idx");
}

const double &SEKParams::get_param(std::size_t index) const
{
    if (index == 0)
        return lengthscale;
    if (index == 1)
        return vertical_lengthscale;
    if (index == 2)
        return noise_variance;
    throw std::invalid_argument("Get Invalid param_idx");
}

AdamParams::AdamParams(double lr, double b1, double b2, double eps, int opt_i) :
    learning_rate(lr), beta1(b1), beta2(b2), epsilon(eps), opt_iter(opt_i)
{ }

std::string AdamParams::repr() const
{
    std::ostringstream oss;
    oss << std::fixed << std::setprecision(8);
    oss << "Hyperparameters: [learning_rate=" << learning_rate << ", beta1=" << beta1 << ", beta2=" << beta2
        << ", epsilon=" << epsilon << ", opt_iter=" << opt_iter << "]";
    return oss.str();
}

}  // namespace gprat_hyper

namespace gprat
{

// ---- targets ----------------------------------------------------------------

CPU::CPU() { }
bool CPU::is_cpu() { return true; }
bool CPU::is_gpu() { return false; }
std::string CPU::repr() const { return "CPU"; }
CPU get_cpu() { return CPU(); }

HIP_GPU::HIP_GPU(int id_, int n_streams_) : id(id_), n_streams(n_streams_)
{
    if (n_streams < 1)
        throw std::invalid_argument("n_streams must be >= 1");
    const int count = gpu_count();
    if (count > 0 && (id < 0 || id >= count))
        throw std::runtime_error("Requested GPU device is not available.");
}

bool HIP_GPU::is_cpu() { return false; }
bool HIP_GPU::is_gpu() { return true; }

std::string HIP_GPU::repr() const
{
    std::ostringstream oss;
    oss << "GPU (HIP): [id=" << id << ", n_streams=" << n_streams << "]";
    return oss.str();
}

gprat_context *HIP_GPU::context()
{
    if (!ctx_)
    {
        gprat_context *c = nullptr;
        check(gprat_context_create(id, n_streams, &c), "GPU target");
        ctx_ = std::shared_ptr<gprat_context>(c, [](gprat_context *p) { gprat_context_destroy(p); });
    }
    return ctx_.get();
}

HIP_GPU get_gpu(int id, int n_streams) { return HIP_GPU(id, n_streams); }
HIP_GPU get_gpu() { return HIP_GPU(0, 1); }

int gpu_count()
{
    int n = 0;
    if (gprat_device_count(&n) != GPRAT_OK)
        return 0;
    return n;
}

void print_available_gpus()
{
    const int n = gpu_count();
    for (int i = 0; i < n; i++)
    {
        char buf[512];
        if (gprat_device_describe(i, buf, sizeof buf) == GPRAT_OK)
            std::cout << buf << std::endl;
    }
}

// ---- data -------------------------------------------------------------------

GP_data::GP_data(const std::string &f_path, int n, int n_reg) : file_path(f_path), n_samples(n), n_regressors(n_reg)
{
    data = utils::load_data(f_path, n, n_reg - 1);
}

// ---- GP ---------------------------------------------------------------------

static gprat_hyper::SEKParams make_params(const std::vector<double> &h)
{
    if (h.size() < 3)
        throw std::invalid_argument("kernel_params needs 3 values (lengthscale, vertical_lengthscale, noise_variance)");
    return gprat_hyper::SEKParams(h[0], h[1], h[2]);
}

static std::vector<bool> check_trainable(std::vector<bool> t)
{
    if (t.size() < 3)
        throw std::invalid_argument("trainable needs 3 flags");
    return t;
}

GP::GP(std::vector<double> input,
       std::vector<double> output,
       int n_tiles,
       int n_tile_size,
       int n_regressors,
       std::vector<double> kernel_hyperparams,
       std::vector<bool> trainable_bool,
       std::shared_ptr<Target> target) :
    training_input_(std::move(input)),
    training_output_(std::move(output)),
    n_tiles_(n_tiles),
    n_tile_size_(n_tile_size),
    trainable_params_(check_trainable(std::move(trainable_bool))),
    target_(std::move(target)),
    n_reg(n_regressors),
    kernel_params(make_params(kernel_hyperparams))
{ }

GP::GP(std::vector<double> input,
       std::vector<double> output,
       int n_tiles,
       int n_tile_size,
       int n_regressors,
       std::vector<double> kernel_hyperparams,
       std::vector<bool> trainable_bool) :
    GP(std::move(input),
       std::move(output),
       n_tiles,
       n_tile_size,
       n_regressors,
       std::move(kernel_hyperparams),
       std::move(trainable_bool),
       std::make_shared<HIP_GPU>(0, 1))
{ }

GP::GP(std::vector<double> input,
       std::vector<double> output,
       int n_tiles,
       int n_tile_size,
       int n_regressors,
       std::vector<double> kernel_hyperparams,
       std::vector<bool> trainable_bool,
       int gpu_id,
       int n_streams) :
    GP(std::move(input),
       std::move(output),
       n_tiles,
       n_tile_size,
       n_regressors,
       std::move(kernel_hyperparams),
       std::move(trainable_bool),
       std::make_shared<HIP_GPU>(gpu_id, n_streams))
{ }

gprat_context *GP::gpu_context()
{
    utils::require_runnable();
    auto gpu = std::dynamic_pointer_cast<HIP_GPU>(target_);
    if (!gpu)
        throw std::runtime_error("gprat_amd executes on the GPU only; construct the GP with a GPU target "
                                 "(gpu_id, n_streams) or without a target");
    return gpu->context();
}

std::string GP::repr() const
{
    std::ostringstream oss;
    oss << std::fixed << std::setprecision(12);
    oss << "Kernel_Params: [lengthscale=" << kernel_params.lengthscale
        << ", vertical_lengthscale=" << kernel_params.vertical_lengthscale
        << ", noise_variance=" << kernel_params.noise_variance << ", n_regressors=" << n_reg
        << "], Trainable_Params: [trainable_params l=" << trainable_params_[0]
        << ", trainable_params v=" << trainable_params_[1] << ", trainable_params n=" << trainable_params_[2]
        << "], Target: [" << target_->repr() << "], n_tiles=" << n_tiles_ << ", n_tile_size=" << n_tile_size_;
    return oss.str();
}

std::vector<double> GP::get_training_input() const { return training_input_; }
std::vector<double> GP::get_training_output() const { return training_output_; }

static std::size_t test_count(int m_tiles, int m_tile_size)
{
    if (m_tiles < 0 || m_tile_size < 0)
        throw std::invalid_argument("m_tiles and m_tile_size must be >= 0");
    return static_cast<std::size_t>(m_tiles) * static_cast<std::size_t>(m_tile_size);
}

std::vector<double> GP::predict(const std::vector<double> &test, int m_tiles, int m_tile_size)
{
    std::vector<double> mean(test_count(m_tiles, m_tile_size));
    const double theta[3] = { kernel_params.lengthscale, kernel_params.vertical_lengthscale, kernel_params.noise_variance };
    check(gprat_predict(gpu_context(), training_input_.data(), training_input_.size(), training_output_.data(),
                        training_output_.size(), test.data(), test.size(), theta, n_tiles_, n_tile_size_, m_tiles,
                        m_tile_size, n_reg, mean.data()),
          "predict");
    return mean;
}

std::vector<std::vector<double>> GP::predict_with_uncertainty(const std::vector<double> &test, int m_tiles, int m_tile_size)
{
    const std::size_t m = test_count(m_tiles, m_tile_size);
    std::vector<double> mean(m), var(m);
    const double theta[3] = { kernel_params.lengthscale, kernel_params.vertical_lengthscale, kernel_params.noise_variance };
    check(gprat_predict_with_uncertainty(gpu_context(), training_input_.data(), training_input_.size(),
                                         training_output_.data(), training_output_.size(), test.data(), test.size(),
                                         theta, n_tiles_, n_tile_size_, m_tiles, m_tile_size, n_reg, mean.data(),
                                         var.data()),
          "predict_with_uncertainty");
    return { std::move(mean), std::move(var) };
}

std::vector<std::vector<double>> GP::predict_with_full_cov(const std::vector<double> &test, int m_tiles, int m_tile_size)
{
    const std::size_t m = test_count(m_tiles, m_tile_size);
    std::vector<double> mean(m), var(m);
    const double theta[3] = { kernel_params.lengthscale, kernel_params.vertical_lengthscale, kernel_params.noise_variance };
    check(gprat_predict_with_full_cov(gpu_context(), training_input_.data(), training_input_.size(),
                                      training_output_.data(), training_output_.size(), test.data(), test.size(), theta,
                                      n_tiles_, n_tile_size_, m_tiles, m_tile_size, n_reg, mean.data(), var.data()),
          "predict_with_full_cov");
    return { std::move(mean), std::move(var) };
}

std::vector<std::vector<double>> GP::predict_full_posterior(const std::vector<double> &test, int m_tiles, int m_tile_size)
{
    const std::size_t m = test_count(m_tiles, m_tile_size);
    std::vector<double> mean(m), cov(m * m);
    const double theta[3] = { kernel_params.lengthscale, kernel_params.vertical_lengthscale, kernel_params.noise_variance };
    check(gprat_predict_full_posterior(gpu_context(), training_input_.data(), training_input_.size(), training_output_.data(),
                                       training_output_.size(), test.data(), test.size(), theta, n_tiles_, n_tile_size_, m_tiles,
                                       m_tile_size, n_reg, mean.data(), cov.data()),
          "predict_full_posterior");
    return { std::move(mean), std::move(cov) };
}

double GP::loss_and_gradient(double grad[3])
{
    const double theta[3] = { kernel_params.lengthscale, kernel_params.vertical_lengthscale, kernel_params.noise_variance };
    const int trainable[3] = { trainable_params_[0] ? 1 : 0, trainable_params_[1] ? 1 : 0, trainable_params_[2] ? 1 : 0 };
    double loss = 0.0;
    check(gprat_loss_gradient(gpu_context(), training_input_.data(), training_input_.size(), training_output_.data(),
                              training_output_.size(), theta, n_tiles_, n_tile_size_, n_reg, trainable, &loss, grad),
          "optimize");
    return loss;
}

std::vector<double> GP::optimize(const gprat_hyper::AdamParams &adam_params)
{
    std::vector<double> losses;
    losses.reserve(static_cast<std::size_t>(std::max(adam_params.opt_iter, 0)));
    for (int iter = 0; iter < adam_params.opt_iter; iter++)
    {
        double grad[3];
        const double loss = loss_and_gradient(grad);
        adam_update(kernel_params, trainable_params_, adam_params, grad, static_cast<std::size_t>(iter));
        losses.push_back(loss);
    }
    return losses;
}

double GP::optimize_step(gprat_hyper::AdamParams &adam_params, int iter)
{
    double grad[3];
    const double loss = loss_and_gradient(grad);
    adam_update(kernel_params, trainable_params_, adam_params, grad, static_cast<std::size_t>(iter));
    return loss;
}

double GP::calculate_loss()
{
    const double theta[3] = { kernel_params.lengthscale, kernel_params.vertical_lengthscale, kernel_params.noise_variance };
    double loss = 0.0;
    check(gprat_compute_loss(gpu_context(), training_input_.data(), training_input_.size(), training_output_.data(),
                             training_output_.size(), theta, n_tiles_, n_tile_size_, n_reg, &loss),
          "calculate_loss");
    return loss;
}

std::vector<std::vector<double>> GP::cholesky()
{
    if (n_tiles_ < 1 || n_tile_size_ < 1)
        throw std::invalid_argument("n_tiles and n_tile_size must be >= 1");
    const std::size_t T = static_cast<std::size_t>(n_tiles_), b = static_cast<std::size_t>(n_tile_size_);
    std::vector<double> packed(T * (T + 1) / 2 * b * b);
    const double theta[3] = { kernel_params.lengthscale, kernel_params.vertical_lengthscale, kernel_params.noise_variance };
    check(gprat_cholesky(gpu_context(), training_input_.data(), training_input_.size(), theta, n_tiles_, n_tile_size_,
                         n_reg, packed.data()),
          "cholesky");
    // T^2 tiles, upper ones empty (core/src/cpu/gp_functions.cpp:35,69-76)
    std::vector<std::vector<double>> result(T * T);
    std::size_t t = 0;
    for (std::size_t i = 0; i < T; i++)
        for (std::size_t j = 0; j <= i; j++, t++)
            result[i * T + j].assign(packed.begin() + static_cast<long>(t * b * b),
                                     packed.begin() + static_cast<long>((t + 1) * b * b));
    return result;
}

}  // namespace gprat

namespace utils
{

// core/src/utils_c.cpp:8-52
int compute_train_tiles(int n_samples, int n_tile_size)
{
    if (n_tile_size > 0)
        return n_samples / n_tile_size;
    throw std::runtime_error("Error: Please specify a valid value for train_tile_size.\n");
}

int compute_train_tile_size(int n_samples, int n_tiles)
{
    if (n_tiles > 0)
        return n_samples / n_tiles;
    throw std::runtime_error("Error: Please specify a valid value for train_tiles.\n");
}

std::pair<int, int> compute_test_tiles(int n_test, int n_tiles, int n_tile_size)
{
    if (n_tile_size <= 0 || n_tiles <= 0)
        throw std::runtime_error("Error: Please specify valid values for n_tiles and n_tile_size.\n");
    if (n_test % n_tile_size > 0)
        return { n_tiles, n_test / n_tiles };
    return { n_test / n_tile_size, n_tile_size };
}

// core/src/utils_c.cpp:54-81: n values, one "%lf" each, after `offset` zeros.
std::vector<double> load_data(const std::string &file_path, int n_samples, int offset)
{
    std::vector<double> data(static_cast<std::size_t>(n_samples + offset), 0.0);
    FILE *f = std::fopen(file_path.c_str(), "r");
    if (f == nullptr)
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

// core/src/utils_c.cpp:83-121
void print_vector(const std::vector<double> &vec, int start, int end, const std::string &separator)
{
    const int size = static_cast<int>(vec.size());
    if (start < 0)
        start += size;
    if (end < 0)
        end += size + 1;
    if (start < 0)
        start = 0;
    if (end > size)
        end = size;
    if (start >= size || start >= end)
    {
        std::cerr << "Invalid range" << std::endl;
        return;
    }
    for (int i = start; i < end; i++)
    {
        std::cout << vec[static_cast<std::size_t>(i)];
        if (i < end - 1)
            std::cout << separator;
    }
    std::cout << std::endl;
}

// The host runtime's life cycle as core/src/utils_c.cpp:123-133 drives HPX:
// start (with "--hpx:threads=N"), suspend, resume, stop, start again. This
// backend needs no host runtime to run the GP path on the GPU, so calls made
// before start or after stop run; a suspended runtime runs no work, so the
// GP API refuses calls until it is resumed.
void start_hpx_runtime(int argc, char **argv)
{
    std::lock_guard<std::mutex> g(g_runtime_mu);
    if (g_runtime_state == RuntimeState::Running || g_runtime_state == RuntimeState::Suspended)
        throw std::runtime_error("the runtime is already started (stop_hpx_runtime first)");
    g_host_threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    // honour "--hpx:threads=N" so callers can query the requested core count
    for (int i = 0; i < argc && argv && argv[i]; i++)
    {
        const std::string a = argv[i];
        const std::string key = "--hpx:threads=";
        if (a.rfind(key, 0) == 0)
            g_host_threads = std::atoi(a.c_str() + key.size());
    }
    g_runtime_state = RuntimeState::Running;
}
void suspend_hpx_runtime()
{
    std::lock_guard<std::mutex> g(g_runtime_mu);
    if (g_runtime_state != RuntimeState::Running)
        throw std::runtime_error("suspend_hpx_runtime: the runtime is not running");
    g_runtime_state = RuntimeState::Suspended;
}
void resume_hpx_runtime()
{
    std::lock_guard<std::mutex> g(g_runtime_mu);
    if (g_runtime_state != RuntimeState::Suspended)
        throw std::runtime_error("resume_hpx_runtime: the runtime is not suspended");
    g_runtime_state = RuntimeState::Running;
}
void stop_hpx_runtime()
{
    std::lock_guard<std::mutex> g(g_runtime_mu);
    if (g_runtime_state != RuntimeState::Running && g_runtime_state != RuntimeState::Suspended)
        throw std::runtime_error("stop_hpx_runtime: the runtime is not started");
    g_runtime_state = RuntimeState::Stopped;
}
int host_threads() { return g_host_threads; }
std::string runtime_state()
{
    std::lock_guard<std::mutex> g(g_runtime_mu);
    switch (g_runtime_state)
    {
    case RuntimeState::Running:
        return "running";
    case RuntimeState::Suspended:
        return "suspended";
    case RuntimeState::Stopped:
        return "stopped";
    default:
        return "not started";
    }
}
void require_runnable()
{
    std::lock_guard<std::mutex> g(g_runtime_mu);
    if (g_runtime_state == RuntimeState::Suspended)
        throw std::runtime_error("the runtime is suspended: resume_hpx_runtime() before calling the GP API");
}

bool compiled_with_cuda() { return true; }
bool compiled_with_hip() { return true; }

}  // namespace utils
