// Host C++ API of the MI355X backend: the same classes and signatures as
// GPRat's public header (core/include/gprat_c.hpp:21-228,
// core/include/gp_kernels.hpp:14-78, core/include/gp_hyperparameters.hpp:12-56,
// core/include/target.hpp:22-204, core/include/utils_c.hpp:19-86), driving the
// gfx950 kernels through the C ABI in include/gprat_amd.h.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

struct gprat_context;

namespace gprat_hyper
{

// core/include/gp_kernels.hpp:14-78
struct SEKParams
{
    double lengthscale;
    double vertical_lengthscale;
    double noise_variance;
    std::vector<double> m_T;  // Adam first moments
    std::vector<double> w_T;  // Adam second moments

    SEKParams(double lengthscale_, double vertical_lengthscale_, double noise_variance_);
    std::size_t size();
    void set_param(std::size_t index, double value);
    const double &get_param(std::size_t index) const;
};

// core/include/gp_hyperparameters.hpp:12-56
struct AdamParams
{
    double learning_rate;
    double beta1;
    double beta2;
    double epsilon;
    int opt_iter;

    AdamParams(double lr = 0.001, double b1 = 0.9, double b2 = 0.999, double eps = 1e-8, int opt_i = 0);
    std::string repr() const;
};

}  // namespace gprat_hyper

namespace gprat
{

// core/include/target.hpp:22-80
struct Target
{
    virtual bool is_cpu() = 0;
    virtual bool is_gpu() = 0;
    virtual std::string repr() const = 0;
    virtual ~Target() { }

  protected:
    Target() = default;
};

struct CPU : public Target
{
    CPU();
    bool is_cpu() override;
    bool is_gpu() override;
    std::string repr() const override;
};

CPU get_cpu();

// The GPU target: one device and a library context created on first use
// (replaces CUDA_GPU, core/include/target.hpp:82-180).
struct HIP_GPU : public Target
{
    int id;
    int n_streams;

    HIP_GPU(int id, int n_streams);
    bool is_cpu() override;
    bool is_gpu() override;
    std::string repr() const override;

    // Library context of this target (created on first call; throws if no
    // device is available).
    gprat_context *context();

  private:
    std::shared_ptr<gprat_context> ctx_;
};

HIP_GPU get_gpu(int id, int n_streams);
HIP_GPU get_gpu();
void print_available_gpus();
int gpu_count();

// core/include/gprat_c.hpp:21-45
struct GP_data
{
    std::string file_path;
    int n_samples;
    int n_regressors;
    std::vector<double> data;
    GP_data(const std::string &file_path, int n, int n_reg);
};

// core/include/gprat_c.hpp:54-228
class GP
{
  private:
    std::vector<double> training_input_;
    std::vector<double> training_output_;
    int n_tiles_;
    int n_tile_size_;
    std::vector<bool> trainable_params_;
    std::shared_ptr<Target> target_;

    gprat_context *gpu_context();
    double loss_and_gradient(double grad[3]);

  public:
    int n_reg;
    gprat_hyper::SEKParams kernel_params;

    GP(std::vector<double> input,
       std::vector<double> output,
       int n_tiles,
       int n_tile_size,
       int n_regressors,
       std::vector<double> kernel_hyperparams,
       std::vector<bool> trainable_bool,
       std::shared_ptr<Target> target);

    // Without an explicit target the GP runs on GPU 0: this library is the
    // GPU backend.
    GP(std::vector<double> input,
       std::vector<double> output,
       int n_tiles,
       int n_tile_size,
       int n_regressors,
       std::vector<double> kernel_hyperparams,
       std::vector<bool> trainable_bool);

    GP(std::vector<double> input,
       std::vector<double> output,
       int n_tiles,
       int n_tile_size,
       int n_regressors,
       std::vector<double> kernel_hyperparams,
       std::vector<bool> trainable_bool,
       int gpu_id,
       int n_streams);

    std::string repr() const;
    std::vector<double> get_training_input() const;
    std::vector<double> get_training_output() const;

    std::vector<double> predict(const std::vector<double> &test_data, int m_tiles, int m_tile_size);
    std::vector<std::vector<double>>
    predict_with_uncertainty(const std::vector<double> &test_data, int m_tiles, int m_tile_size);
    std::vector<std::vector<double>>
    predict_with_full_cov(const std::vector<double> &test_data, int m_tiles, int m_tile_size);
    // {mean, the whole m x m latent posterior covariance, row-major}: the
    // matrix the reference's predict_with_full_cov forms and reduces to its
    // diagonal (core/src/cpu/gp_functions.cpp:452-696).
    std::vector<std::vector<double>>
    predict_full_posterior(const std::vector<double> &test_data, int m_tiles, int m_tile_size);
    std::vector<double> optimize(const gprat_hyper::AdamParams &adam_params);
    double optimize_step(gprat_hyper::AdamParams &adam_params, int iter);
    double calculate_loss();
    std::vector<std::vector<double>> cholesky();
};

}  // namespace gprat

namespace utils
{

// core/include/utils_c.hpp:19-86
int compute_train_tiles(int n_samples, int n_tile_size);
int compute_train_tile_size(int n_samples, int n_tiles);
std::pair<int, int> compute_test_tiles(int n_test, int n_tiles, int n_tile_size);
std::vector<double> load_data(const std::string &file_path, int n_samples, int offset);
void print_vector(const std::vector<double> &vec, int start, int end, const std::string &separator);
// The reference's runtime life cycle (core/src/utils_c.cpp:123-133): start
// (records "--hpx:threads=N"), suspend, resume, stop, and start again; out
// of order calls throw std::runtime_error as HPX refuses them. The GPU path
// needs no host runtime: the GP API runs before start and after stop, and
// refuses calls while the runtime is suspended.
void start_hpx_runtime(int argc, char **argv);
void resume_hpx_runtime();
void suspend_hpx_runtime();
void stop_hpx_runtime();
int host_threads();
// "not started", "running", "suspended" or "stopped"
std::string runtime_state();
// Throws if the runtime is suspended.
void require_runnable();
// True when the GPU backend is compiled in (always, for this library).
bool compiled_with_cuda();
bool compiled_with_hip();

}  // namespace utils
