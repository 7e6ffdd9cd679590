// TEST INFRASTRUCTURE — the parity oracle. Nothing in the product (gprat_amd/)
// may include, link or call this; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker.
//
// A CPU restatement of GPRat's `cpu::` GP path (core/src/cpu/*.cpp), executed
// as a tile dataflow graph on a host thread pool that stands in for HPX, with
// the reference's exact per-tile CBLAS/LAPACKE calls bound at run time to
// oneMKL (sequential) or, failing that, to numpy's bundled OpenBLAS.
// Each function cites the reference file:line it restates.
#pragma once

#include <cstddef>
#include <string>
#include <utility>
#include <vector>

namespace oracle
{

// core/include/gp_kernels.hpp:14-78, core/src/gp_kernels.cpp:7-53
struct SEKParams
{
    double lengthscale;
    double vertical_lengthscale;
    double noise_variance;
    std::vector<double> m_T;
    std::vector<double> w_T;

    SEKParams(double l, double v, double n);
    std::size_t size() const { return 3; }
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
    AdamParams(double lr = 0.001, double b1 = 0.9, double b2 = 0.999, double eps = 1e-8, int it = 0) :
        learning_rate(lr), beta1(b1), beta2(b2), epsilon(eps), opt_iter(it)
    { }
};

// core/src/utils_c.cpp:8-81
int compute_train_tiles(int n_samples, int n_tile_size);
int compute_train_tile_size(int n_samples, int n_tiles);
std::pair<int, int> compute_test_tiles(int n_test, int n_tiles, int n_tile_size);
std::vector<double> load_data(const std::string &file_path, int n_samples, int offset);

// Host pool (the HPX worker count, `--hpx:threads`).
void set_threads(unsigned n);
unsigned threads();
// Which BLAS the tile calls are bound to ("mkl", "openblas64").
std::string blas_name();

using Matrix = std::vector<std::vector<double>>;

// core/src/cpu/gp_functions.cpp:17-78
Matrix cholesky(const std::vector<double> &training_input,
                const SEKParams &sek,
                int n_tiles,
                int n_tile_size,
                int n_regressors);

// core/src/cpu/gp_functions.cpp:80-216
std::vector<double> predict(const std::vector<double> &training_input,
                            const std::vector<double> &training_output,
                            const std::vector<double> &test_input,
                            const SEKParams &sek,
                            int n_tiles,
                            int n_tile_size,
                            int m_tiles,
                            int m_tile_size,
                            int n_regressors);

// core/src/cpu/gp_functions.cpp:218-450
Matrix predict_with_uncertainty(const std::vector<double> &training_input,
                                const std::vector<double> &training_output,
                                const std::vector<double> &test_input,
                                const SEKParams &sek,
                                int n_tiles,
                                int n_tile_size,
                                int m_tiles,
                                int m_tile_size,
                                int n_regressors);

// core/src/cpu/gp_functions.cpp:452-696
Matrix predict_with_full_cov(const std::vector<double> &training_input,
                             const std::vector<double> &training_output,
                             const std::vector<double> &test_input,
                             const SEKParams &sek,
                             int n_tiles,
                             int n_tile_size,
                             int m_tiles,
                             int m_tile_size,
                             int n_regressors);

// core/src/cpu/gp_functions.cpp:700-781
double compute_loss(const std::vector<double> &training_input,
                    const std::vector<double> &training_output,
                    const SEKParams &sek,
                    int n_tiles,
                    int n_tile_size,
                    int n_regressors);

// core/src/cpu/gp_functions.cpp:783-1030
std::vector<double> optimize(const std::vector<double> &training_input,
                             const std::vector<double> &training_output,
                             int n_tiles,
                             int n_tile_size,
                             int n_regressors,
                             const AdamParams &adam,
                             SEKParams &sek,
                             const std::vector<bool> &trainable);

// core/src/cpu/gp_functions.cpp:1032-1269
double optimize_step(const std::vector<double> &training_input,
                     const std::vector<double> &training_output,
                     int n_tiles,
                     int n_tile_size,
                     int n_regressors,
                     AdamParams &adam,
                     SEKParams &sek,
                     const std::vector<bool> &trainable,
                     int iter);

// FP32 comparator of the mixed-precision path: predict_with_uncertainty's
// DAG with a reduced-precision Cholesky, level 1 = SYRK/GEMM trailing
// products through the FP32 adapter (core/src/cpu/adapter_cblas_fp32.cpp:54-100)
// subtracted from FP64 tiles, level 2 = every factor op through the FP32
// adapter (:14-100); alpha refined in FP64 against K until the correction
// is at most 1e-12 max|alpha|; the variance solve, against the reduced
// factor, is not refined; the loss uses the reduced factor's diagonal.
// Level 0 is the FP64 path plus the refinement.
struct Reduced
{
    std::vector<double> mean, var;
    double loss = 0.0;
    int refine_steps = 0;
};
Reduced predict_with_uncertainty_reduced(const std::vector<double> &training_input,
                                         const std::vector<double> &training_output,
                                         const std::vector<double> &test_input,
                                         const SEKParams &sek,
                                         int n_tiles,
                                         int n_tile_size,
                                         int m_tiles,
                                         int m_tile_size,
                                         int n_regressors,
                                         int level);

// Scalar pieces exposed for closed-form tests (core/src/cpu/gp_optimizer.cpp:12-36).
double to_constrained(double parameter, bool noise);
double to_unconstrained(double parameter, bool noise);
double compute_sigmoid(double parameter);

}  // namespace oracle
