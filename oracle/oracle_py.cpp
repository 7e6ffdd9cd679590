// TEST INFRASTRUCTURE — Python face of the parity oracle (module `gprat_oracle`).
// Mirrors the reference's pybind module `gprat` (bindings/gprat_py.cpp:10-142,
// bindings/utils_py.cpp:43-109) for the CPU target, so tests can call the
// oracle and the product with the same code. Adds `cholesky()` and read access
// to the hyperparameters and Adam moments, which tests need.
#include "gprat_oracle.hpp"

#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>

namespace py = pybind11;

namespace
{

struct Data
{
    std::string file_path;
    int n_samples;
    int n_regressors;
    std::vector<double> data;
    Data(const std::string &f, int n, int r) :
        file_path(f), n_samples(n), n_regressors(r), data(oracle::load_data(f, n, r - 1))
    { }
};

// core/include/gprat_c.hpp:54-228 (CPU target only)
class GP
{
  public:
    GP(std::vector<double> in,
       std::vector<double> out,
       int n_tiles,
       int n_tile_size,
       int n_reg,
       std::vector<double> theta,
       std::vector<bool> trainable) :
        in_(std::move(in)),
        out_(std::move(out)),
        n_tiles_(n_tiles),
        n_tile_size_(n_tile_size),
        trainable_(std::move(trainable)),
        n_reg(n_reg),
        sek(theta.at(0), theta.at(1), theta.at(2))
    { }

    std::vector<double> predict(const std::vector<double> &z, int mt, int ms)
    {
        return oracle::predict(in_, out_, z, sek, n_tiles_, n_tile_size_, mt, ms, n_reg);
    }
    oracle::Matrix predict_with_uncertainty(const std::vector<double> &z, int mt, int ms)
    {
        return oracle::predict_with_uncertainty(in_, out_, z, sek, n_tiles_, n_tile_size_, mt, ms, n_reg);
    }
    oracle::Matrix predict_with_full_cov(const std::vector<double> &z, int mt, int ms)
    {
        return oracle::predict_with_full_cov(in_, out_, z, sek, n_tiles_, n_tile_size_, mt, ms, n_reg);
    }
    std::vector<double> optimize(const oracle::AdamParams &a)
    {
        return oracle::optimize(in_, out_, n_tiles_, n_tile_size_, n_reg, a, sek, trainable_);
    }
    double optimize_step(oracle::AdamParams &a, int iter)
    {
        return oracle::optimize_step(in_, out_, n_tiles_, n_tile_size_, n_reg, a, sek, trainable_, iter);
    }
    double calculate_loss() { return oracle::compute_loss(in_, out_, sek, n_tiles_, n_tile_size_, n_reg); }
    oracle::Reduced predict_reduced(const std::vector<double> &z, int mt, int ms, int level)
    {
        return oracle::predict_with_uncertainty_reduced(in_, out_, z, sek, n_tiles_, n_tile_size_, mt, ms, n_reg, level);
    }
    oracle::Matrix cholesky() { return oracle::cholesky(in_, sek, n_tiles_, n_tile_size_, n_reg); }

    std::vector<double> in_, out_;
    int n_tiles_, n_tile_size_;
    std::vector<bool> trainable_;
    int n_reg;
    oracle::SEKParams sek;
};

}  // namespace

PYBIND11_MODULE(gprat_oracle, m)
{
    m.doc() = "CPU parity oracle: restatement of GPRat's cpu:: GP path (test infrastructure)";

    py::class_<Data>(m, "GP_data")
        .def(py::init<std::string, int, int>(), py::arg("file_path"), py::arg("n_samples"), py::arg("n_regressors"))
        .def_readonly("n_samples", &Data::n_samples)
        .def_readonly("n_regressors", &Data::n_regressors)
        .def_readonly("file_path", &Data::file_path)
        .def_readonly("data", &Data::data);

    py::class_<oracle::AdamParams>(m, "AdamParams")
        .def(py::init<double, double, double, double, int>(),
             py::arg("learning_rate") = 0.001,
             py::arg("beta1") = 0.9,
             py::arg("beta2") = 0.999,
             py::arg("epsilon") = 1e-8,
             py::arg("opt_iter") = 0)
        .def_readwrite("learning_rate", &oracle::AdamParams::learning_rate)
        .def_readwrite("beta1", &oracle::AdamParams::beta1)
        .def_readwrite("beta2", &oracle::AdamParams::beta2)
        .def_readwrite("epsilon", &oracle::AdamParams::epsilon)
        .def_readwrite("opt_iter", &oracle::AdamParams::opt_iter);

    py::class_<GP>(m, "GP")
        .def(py::init<std::vector<double>, std::vector<double>, int, int, int, std::vector<double>, std::vector<bool>>(),
             py::arg("input_data"),
             py::arg("output_data"),
             py::arg("n_tiles"),
             py::arg("n_tile_size"),
             py::arg("n_reg") = 8,
             py::arg("kernel_params") = std::vector<double>{ 1.0, 1.0, 0.1 },
             py::arg("trainable") = std::vector<bool>{ true, true, true })
        .def_readwrite("n_reg", &GP::n_reg)
        .def("predict", &GP::predict, py::arg("test_data"), py::arg("m_tiles"), py::arg("m_tile_size"))
        .def("predict_with_uncertainty",
             &GP::predict_with_uncertainty,
             py::arg("test_data"),
             py::arg("m_tiles"),
             py::arg("m_tile_size"))
        .def("predict_with_full_cov", &GP::predict_with_full_cov, py::arg("test_data"), py::arg("m_tiles"), py::arg("m_tile_size"))
        .def("optimize", &GP::optimize, py::arg("AdamParams"))
        .def("optimize_step", &GP::optimize_step, py::arg("AdamParams"), py::arg("iter"))
        .def("compute_loss", &GP::calculate_loss)
        .def(
            "predict_with_uncertainty_reduced",
            [](GP &g, const std::vector<double> &z, int mt, int ms, int level)
            {
                oracle::Reduced r = g.predict_reduced(z, mt, ms, level);
                return py::make_tuple(r.mean, r.var, r.loss, r.refine_steps);
            },
            py::arg("test_data"),
            py::arg("m_tiles"),
            py::arg("m_tile_size"),
            py::arg("level"))
        .def("cholesky", &GP::cholesky)
        .def_property_readonly("hyperparameters",
                               [](const GP &g)
                               {
                                   return std::vector<double>{ g.sek.lengthscale, g.sek.vertical_lengthscale,
                                                               g.sek.noise_variance };
                               })
        .def_property_readonly("adam_moments", [](const GP &g) { return std::make_pair(g.sek.m_T, g.sek.w_T); });

    m.def("compute_train_tiles", &oracle::compute_train_tiles, py::arg("n_samples"), py::arg("n_tile_size"));
    m.def("compute_train_tile_size", &oracle::compute_train_tile_size, py::arg("n_samples"), py::arg("n_tiles"));
    m.def("compute_test_tiles", &oracle::compute_test_tiles, py::arg("m_samples"), py::arg("n_tiles"), py::arg("n_tile_size"));
    m.def(
        "start_hpx",
        [](std::vector<std::string>, std::size_t n_cores) { oracle::set_threads(static_cast<unsigned>(n_cores)); },
        py::arg("args"),
        py::arg("n_cores"));
    m.def("stop_hpx", []() { });
    m.def("set_threads", [](unsigned n) { oracle::set_threads(n); });
    m.def("threads", &oracle::threads);
    m.def("blas_name", &oracle::blas_name);
    m.def("to_constrained", &oracle::to_constrained);
    m.def("to_unconstrained", &oracle::to_unconstrained);
    m.def("compute_sigmoid", &oracle::compute_sigmoid);
}
