// pybind11 module `gprat` of the MI355X backend: the same Python surface as
// the reference module (bindings/gprat_py.cpp:10-142, bindings/utils_py.cpp:16-109,
// bindings/module.cpp:9-22), plus SEKParams (which the reference exposes as
// GP.kernel_params but never registers), GP.cholesky, and a resident Model
// for fit-once / predict-many use.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>

#include "gprat.hpp"
#include "gprat_amd.h"

namespace py = pybind11;

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

// A fitted GP resident in HBM (factor and alpha stay on the device).
class Model
{
  public:
    Model(std::shared_ptr<gprat::HIP_GPU> gpu,
          const std::vector<double> &input,
          const std::vector<double> &output,
          int n,
          int n_reg,
          std::vector<double> theta,
          int precision) :
        gpu_(std::move(gpu)), n_(n), n_reg_(n_reg)
    {
        if (theta.size() < 3)
            throw std::invalid_argument("kernel_params needs 3 values");
        gprat_model *m = nullptr;
        {
            py::gil_scoped_release nogil;
            check(gprat_model_fit(gpu_->context(), input.data(), input.size(), output.data(), output.size(), n, n_reg,
                                  theta.data(), precision, &m),
                  "Model");
        }
        m_ = m;
    }
    ~Model()
    {
        if (m_)
            gprat_model_release(m_);
    }
    Model(const Model &) = delete;
    Model &operator=(const Model &) = delete;

    std::vector<std::vector<double>> predict(const std::vector<double> &test, int m, bool with_variance)
    {
        if (m < 0)
            throw std::invalid_argument("m must be >= 0");
        std::vector<double> mean(static_cast<std::size_t>(m)), var(with_variance ? static_cast<std::size_t>(m) : 0);
        {
            py::gil_scoped_release nogil;
            check(gprat_model_predict(m_, test.data(), test.size(), m, mean.data(), with_variance ? var.data() : nullptr),
                  "Model.predict");
        }
        if (with_variance)
            return { std::move(mean), std::move(var) };
        return { std::move(mean) };
    }
    double loss()
    {
        double l = 0.0;
        check(gprat_model_loss(m_, &l), "Model.loss");
        return l;
    }
    std::size_t bytes() const
    {
        std::size_t b = 0;
        check(gprat_model_bytes(m_, &b), "Model.bytes");
        return b;
    }
    int n() const { return n_; }
    int n_reg() const { return n_reg_; }

  private:
    std::shared_ptr<gprat::HIP_GPU> gpu_;
    gprat_model *m_ = nullptr;
    int n_, n_reg_;
};

void start_hpx_wrapper(std::vector<std::string> args, std::size_t n_cores)
{
    if (args.empty())
        args.push_back("gprat");
    args.push_back("--hpx:threads=" + std::to_string(n_cores));
    std::vector<char *> argv;
    for (auto &a : args)
        argv.push_back(&a[0]);
    argv.push_back(nullptr);
    utils::start_hpx_runtime(static_cast<int>(args.size()), argv.data());
}

}  // namespace

PYBIND11_MODULE(gprat, m)
{
    m.doc() = "GPRat on MI355X: exact GP regression with gfx950 HIP kernels";

    py::class_<gprat::GP_data>(m, "GP_data", "Class representing Gaussian Process data.")
        .def(py::init<std::string, int, int>(), py::arg("file_path"), py::arg("n_samples"), py::arg("n_regressors"))
        .def_readonly("n_samples", &gprat::GP_data::n_samples)
        .def_readonly("n_regressors", &gprat::GP_data::n_regressors)
        .def_readonly("file_path", &gprat::GP_data::file_path)
        .def_readonly("data", &gprat::GP_data::data);

    py::class_<gprat_hyper::AdamParams>(m, "AdamParams")
        .def(py::init<double, double, double, double, int>(),
             py::arg("learning_rate") = 0.001,
             py::arg("beta1") = 0.9,
             py::arg("beta2") = 0.999,
             py::arg("epsilon") = 1e-8,
             py::arg("opt_iter") = 0)
        .def_readwrite("learning_rate", &gprat_hyper::AdamParams::learning_rate)
        .def_readwrite("beta1", &gprat_hyper::AdamParams::beta1)
        .def_readwrite("beta2", &gprat_hyper::AdamParams::beta2)
        .def_readwrite("epsilon", &gprat_hyper::AdamParams::epsilon)
        .def_readwrite("opt_iter", &gprat_hyper::AdamParams::opt_iter)
        .def("__repr__", &gprat_hyper::AdamParams::repr);

    py::class_<gprat_hyper::SEKParams>(m, "SEKParams")
        .def(py::init<double, double, double>(),
             py::arg("lengthscale"),
             py::arg("vertical_lengthscale"),
             py::arg("noise_variance"))
        .def_readwrite("lengthscale", &gprat_hyper::SEKParams::lengthscale)
        .def_readwrite("vertical_lengthscale", &gprat_hyper::SEKParams::vertical_lengthscale)
        .def_readwrite("noise_variance", &gprat_hyper::SEKParams::noise_variance)
        .def_readwrite("m_T", &gprat_hyper::SEKParams::m_T)
        .def_readwrite("w_T", &gprat_hyper::SEKParams::w_T);

    py::class_<gprat::GP>(m, "GP")
        .def(py::init<std::vector<double>, std::vector<double>, int, int, int, std::vector<double>, std::vector<bool>>(),
             py::arg("input_data"),
             py::arg("output_data"),
             py::arg("n_tiles"),
             py::arg("n_tile_size"),
             py::arg("n_reg") = 8,
             py::arg("kernel_params") = std::vector<double>{ 1.0, 1.0, 0.1 },
             py::arg("trainable") = std::vector<bool>{ true, true, true })
        .def(py::init<std::vector<double>, std::vector<double>, int, int, int, std::vector<double>, std::vector<bool>, int, int>(),
             py::arg("input_data"),
             py::arg("output_data"),
             py::arg("n_tiles"),
             py::arg("n_tile_size"),
             py::arg("n_reg") = 8,
             py::arg("kernel_params") = std::vector<double>{ 1.0, 1.0, 0.1 },
             py::arg("trainable") = std::vector<bool>{ true, true, true },
             py::arg("gpu_id") = 0,
             py::arg("n_streams") = 1)
        .def_readwrite("n_reg", &gprat::GP::n_reg)
        .def_readwrite("kernel_params", &gprat::GP::kernel_params)
        .def("__repr__", &gprat::GP::repr)
        .def("get_input_data", &gprat::GP::get_training_input)
        .def("get_output_data", &gprat::GP::get_training_output)
        .def("predict", &gprat::GP::predict, py::arg("test_data"), py::arg("m_tiles"), py::arg("m_tile_size"),
             py::call_guard<py::gil_scoped_release>())
        .def("predict_with_uncertainty", &gprat::GP::predict_with_uncertainty, py::arg("test_data"), py::arg("m_tiles"),
             py::arg("m_tile_size"), py::call_guard<py::gil_scoped_release>())
        .def("predict_with_full_cov", &gprat::GP::predict_with_full_cov, py::arg("test_data"), py::arg("m_tiles"),
             py::arg("m_tile_size"), py::call_guard<py::gil_scoped_release>())
        .def("predict_full_posterior", &gprat::GP::predict_full_posterior, py::arg("test_data"), py::arg("m_tiles"),
             py::arg("m_tile_size"), py::call_guard<py::gil_scoped_release>())
        .def("optimize", &gprat::GP::optimize, py::arg("AdamParams"), py::call_guard<py::gil_scoped_release>())
        .def("optimize_step", &gprat::GP::optimize_step, py::arg("AdamParams"), py::arg("iter"),
             py::call_guard<py::gil_scoped_release>())
        .def("compute_loss", &gprat::GP::calculate_loss, py::call_guard<py::gil_scoped_release>())
        .def("cholesky", &gprat::GP::cholesky, py::call_guard<py::gil_scoped_release>());

    py::class_<gprat::HIP_GPU, std::shared_ptr<gprat::HIP_GPU>>(m, "HIP_GPU")
        .def(py::init<int, int>(), py::arg("gpu_id") = 0, py::arg("n_streams") = 1)
        .def_readonly("id", &gprat::HIP_GPU::id)
        .def_readonly("n_streams", &gprat::HIP_GPU::n_streams)
        .def("__repr__", &gprat::HIP_GPU::repr);

    py::class_<Model>(m, "Model", "A fitted GP whose Cholesky factor stays resident on the GPU.")
        .def(py::init<std::shared_ptr<gprat::HIP_GPU>, const std::vector<double> &, const std::vector<double> &, int, int,
                      std::vector<double>, int>(),
             py::arg("gpu"),
             py::arg("input_data"),
             py::arg("output_data"),
             py::arg("n"),
             py::arg("n_reg") = 8,
             py::arg("kernel_params") = std::vector<double>{ 1.0, 1.0, 0.1 },
             py::arg("precision") = static_cast<int>(GPRAT_PRECISION_FP64))
        .def("predict", &Model::predict, py::arg("test_data"), py::arg("m"), py::arg("with_variance") = true)
        .def("loss", &Model::loss)
        .def_property_readonly("bytes", &Model::bytes)
        .def_property_readonly("n", &Model::n)
        .def_property_readonly("n_reg", &Model::n_reg);

    m.attr("PRECISION_FP64") = static_cast<int>(GPRAT_PRECISION_FP64);
    m.attr("PRECISION_MIXED") = static_cast<int>(GPRAT_PRECISION_MIXED);

    m.def("compute_train_tiles", &utils::compute_train_tiles, py::arg("n_samples"), py::arg("n_tile_size"));
    m.def("compute_train_tile_size", &utils::compute_train_tile_size, py::arg("n_samples"), py::arg("n_tiles"));
    m.def("compute_test_tiles", &utils::compute_test_tiles, py::arg("m_samples"), py::arg("n_tiles"), py::arg("n_tile_size"));
    m.def("print_vector", &utils::print_vector, py::arg("vec"), py::arg("start") = 0, py::arg("end") = -1,
          py::arg("separator") = " ");
    m.def("start_hpx", &start_hpx_wrapper, py::arg("args"), py::arg("n_cores"));
    m.def("resume_hpx", &utils::resume_hpx_runtime);
    m.def("runtime_state", &utils::runtime_state);
    m.def("host_threads", &utils::host_threads);
    m.def("suspend_hpx", &utils::suspend_hpx_runtime);
    m.def("stop_hpx", &utils::stop_hpx_runtime);
    m.def("compiled_with_cuda", &utils::compiled_with_cuda, "True: the GPU backend is compiled in");
    m.def("compiled_with_hip", &utils::compiled_with_hip);
    m.def("print_available_gpus", &gprat::print_available_gpus);
    m.def("gpu_count", &gprat::gpu_count);
    m.def("version", []() { return std::string(gprat_version()); });
}
