// pybind11 module `_kpmx`: the binding of the KPM engine, mirroring the reference binding
// (cppmodule/src/kpm.cpp:8-126: the `kpm` factory, KPM, KPMStats, KPMKernel and the kernels).
// The Hamiltonian arrives as CSR arrays (the model builder lives in Python); the compute backend
// is always the GPU engine.
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "kpm_coret.hpp"

namespace py = pybind11;
using namespace kpmx_host;

namespace {

template <class T>
Csr<T> csr_from(py::tuple shape, py::array indptr, py::array indices, py::array data) {
    Csr<T> m;
    m.rows = shape[0].cast<idx_t>();
    m.cols = shape[1].cast<idx_t>();
    auto ip = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(indptr);
    auto ix = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(indices);
    auto dv = py::array_t<T, py::array::c_style | py::array::forcecast>(data);
    if (ip.size() != m.rows + 1) throw std::invalid_argument("CSR indptr has the wrong size");
    if (ix.size() != dv.size()) throw std::invalid_argument("CSR indices and data differ in size");
    m.indptr.assign(ip.data(), ip.data() + ip.size());
    m.indices.assign(ix.data(), ix.data() + ix.size());
    m.data.assign(dv.data(), dv.data() + dv.size());
    return m;
}

Csr<cd> csr_cd(py::object op) {
    Csr<cd> m;
    if (op.is_none()) return m;
    py::tuple shape = op.attr("shape");
    if (shape[0].cast<idx_t>() == 0 || py::len(op.attr("data")) == 0) {
        m.rows = 0;
        return m;
    }
    return csr_from<cd>(shape, op.attr("indptr"), op.attr("indices"), op.attr("data"));
}

struct KPMCore {
    std::unique_ptr<CoreBase> core;
    std::string dtype;
};

py::array_t<double> as_array(std::vector<double> const& v) { return py::array_t<double>(v.size(), v.data()); }

// numpy array -> std::vector in one copy (a std::vector parameter would convert element by element)
template <class T>
using carray = py::array_t<T, py::array::c_style | py::array::forcecast>;
template <class T>
std::vector<T> to_vec(carray<T> const& a) {
    return std::vector<T>(a.data(), a.data() + a.size());
}
py::array_t<cd> as_array(std::vector<cd> const& v) { return py::array_t<cd>(v.size(), v.data()); }

}  // namespace

PYBIND11_MODULE(_kpmx, m) {
    m.doc() = "MI355X KPM engine: host core and HIP compute backend";

    py::class_<Kernel>(m, "KPMKernel")
        .def_readonly("damping_coefficients", &Kernel::damping_coefficients)
        .def_readonly("required_num_moments", &Kernel::required_num_moments);
    m.def("jackson_kernel", &jackson_kernel);
    m.def("lorentz_kernel", &lorentz_kernel, py::arg("lambda_value") = 4.0);
    m.def("dirichlet_kernel", &dirichlet_kernel);
    m.def("round_num_moments", &round_num_moments);

    py::class_<Stats>(m, "KPMStats")
        .def_readonly("num_moments", &Stats::num_moments)
        .def_readonly("uses_full_system", &Stats::uses_full_system)
        .def_readonly("nnz", &Stats::nnz)
        .def_readonly("opt_nnz", &Stats::opt_nnz)
        .def_readonly("vec", &Stats::vec)
        .def_readonly("opt_vec", &Stats::opt_vec)
        .def_readonly("matrix_memory", &Stats::matrix_memory)
        .def_readonly("vector_memory", &Stats::vector_memory)
        .def_readonly("multiplier", &Stats::multiplier)
        .def_property_readonly("eps", &Stats::eps)
        .def("ops", &Stats::ops)
        .def_property_readonly("hamiltonian_time", [](Stats const& s) { return s.hamiltonian_timer.elapsed; })
        .def_property_readonly("moments_time", [](Stats const& s) { return s.moments_timer.elapsed; });

    py::class_<KPMCore>(m, "KPM")
        .def_property_readonly("scaling_factors",
                               [](KPMCore& k) {
                                   auto s = k.core->scaling_factors();
                                   return py::make_tuple(s.a, s.b);
                               })
        .def_property_readonly("energy_range",
                               [](KPMCore& k) { return py::make_tuple(k.core->min_energy(), k.core->max_energy()); })
        .def_property_readonly("lanczos_loops", [](KPMCore& k) { return k.core->lanczos_loops(); })
        .def_property_readonly("kernel", [](KPMCore& k) { return k.core->config().kernel; })
        .def_property_readonly("stats", [](KPMCore& k) { return k.core->stats(); })
        .def_property_readonly("dtype", [](KPMCore& k) { return k.dtype; })
        .def_property_readonly("size", [](KPMCore& k) { return k.core->size(); })
        .def_property_readonly("is_complex", [](KPMCore& k) { return k.core->is_complex(); })
        .def("set_locality_order",
             [](KPMCore& k, carray<int32_t> order) { k.core->set_locality_order(to_vec(order)); })
        .def("moments",
             [](KPMCore& k, idx_t num_moments, carray<cd> alpha, carray<cd> beta, py::object op) {
                 return as_array(k.core->moments(num_moments, to_vec(alpha), to_vec(beta), csr_cd(op)));
             })
        .def("calc_ldos_indices",
             [](KPMCore& k, std::vector<double> energy, double broadening, std::vector<idx_t> idx) {
                 idx_t cols = 0;
                 std::vector<double> r;
                 {
                     py::gil_scoped_release nogil;
                     r = k.core->ldos(idx, energy, broadening, &cols);
                 }
                 py::array_t<double> out({(py::ssize_t)energy.size(), (py::ssize_t)cols});
                 std::memcpy(out.mutable_data(), r.data(), r.size() * sizeof(double));
                 return out;
             })
        .def("calc_dos",
             [](KPMCore& k, std::vector<double> energy, double broadening, idx_t num_random) {
                 std::vector<double> r;
                 {
                     py::gil_scoped_release nogil;
                     r = k.core->dos(energy, broadening, num_random);
                 }
                 return as_array(r);
             })
        .def("dos_moments",
             [](KPMCore& k, idx_t num_moments, idx_t num_random) {
                 std::vector<cd> r;
                 {
                     py::gil_scoped_release nogil;
                     r = k.core->dos_moments(num_moments, num_random);
                 }
                 return as_array(r);
             })
        .def("calc_greens_vector",
             [](KPMCore& k, idx_t row, std::vector<idx_t> cols, std::vector<double> energy, double broadening) {
                 std::vector<std::vector<cd>> r;
                 {
                     py::gil_scoped_release nogil;
                     r = k.core->greens_vector(row, cols, energy, broadening);
                 }
                 py::list out;
                 for (auto& g : r) out.append(as_array(g));
                 return out;
             })
        .def("calc_conductivity",
             [](KPMCore& k, std::vector<double> mu, double broadening, double temperature, carray<float> left_a,
                carray<float> right_a, idx_t num_random, idx_t num_points) {
                 std::vector<cd> r;
                 const std::vector<float> left = to_vec(left_a), right = to_vec(right_a);
                 {
                     py::gil_scoped_release nogil;
                     r = k.core->conductivity(left, right, mu, broadening, temperature, num_random, num_points);
                 }
                 return as_array(r);
             });

    m.def(
        "kpm",
        [](py::tuple shape, py::array indptr, py::array indices, py::array data, std::pair<float, float> energy,
           py::object kernel, std::string matrix_format, bool optimal_size, bool interleaved, float lanczos_precision,
           py::object progress_callback, int device, std::string kubo_precision, bool autotune) {
            Config cfg;
            cfg.autotune = autotune;
            cfg.min_energy = energy.first;
            cfg.max_energy = energy.second;
            if (!kernel.is_none()) cfg.kernel = kernel.cast<Kernel>();
            cfg.matrix_format = matrix_format == "ELL" ? MatrixFormat::ELL : MatrixFormat::CSR;
            cfg.algorithm.optimal_size = optimal_size;
            cfg.algorithm.interleaved = interleaved;
            cfg.lanczos_precision = lanczos_precision;
            cfg.device = device;
            ProgressCallback progress;
            if (!progress_callback.is_none()) {
                auto fn = std::make_shared<py::object>(progress_callback);
                progress = [fn](idx_t delta, idx_t total) {
                    py::gil_scoped_acquire gil;
                    (*fn)(delta, total);
                };
            }
            int kubo_mode = KPMX_KUBO_FP64;
            if (kubo_precision == "fp32_split") kubo_mode = KPMX_KUBO_FP32_SPLIT;
            else if (kubo_precision == "bf16x3") kubo_mode = KPMX_KUBO_BF16X3;
            else if (kubo_precision != "fp64")
                throw std::invalid_argument("kubo_precision must be 'fp64', 'fp32_split' or 'bf16x3'");
            auto dt = py::dtype::from_args(data.attr("dtype"));
            const char kind = dt.kind();
            const auto itemsize = dt.itemsize();
            KPMCore k;
            if (kind == 'f' && itemsize == 4) {
                k.core = std::make_unique<CoreT<float>>(csr_from<float>(shape, indptr, indices, data), cfg, progress, kubo_mode);
                k.dtype = "float32";
            } else if (kind == 'f' && itemsize == 8) {
                k.core = std::make_unique<CoreT<double>>(csr_from<double>(shape, indptr, indices, data), cfg, progress, kubo_mode);
                k.dtype = "float64";
            } else if (kind == 'c' && itemsize == 8) {
                k.core = std::make_unique<CoreT<std::complex<float>>>(csr_from<std::complex<float>>(shape, indptr, indices, data),
                                                                      cfg, progress, kubo_mode);
                k.dtype = "complex64";
            } else if (kind == 'c' && itemsize == 16) {
                k.core = std::make_unique<CoreT<cd>>(csr_from<cd>(shape, indptr, indices, data), cfg, progress, kubo_mode);
                k.dtype = "complex128";
            } else {
                throw std::invalid_argument("unsupported Hamiltonian dtype");
            }
            return k;
        },
        py::arg("shape"), py::arg("indptr"), py::arg("indices"), py::arg("data"),
        py::arg("energy_range") = std::make_pair(0.0f, 0.0f), py::arg("kernel") = py::none(),
        py::arg("matrix_format") = "ELL", py::arg("optimal_size") = true, py::arg("interleaved") = true,
        py::arg("lanczos_precision") = 0.002f, py::arg("progress_callback") = py::none(), py::arg("device") = 0,
        py::arg("kubo_precision") = "fp64", py::arg("autotune") = false);
    // The optimized Hamiltonian the engine uploads: H2 = (H - b) 2/a, optionally BFS-reordered
    // from index `src` (OptimizedHamiltonian::optimize_for), with the reorder map and slice borders.
    m.def(
        "optimize",
        [](py::tuple shape, py::array indptr, py::array indices, py::array data, std::pair<float, float> energy,
           bool reorder, std::vector<idx_t> src, std::vector<idx_t> dest) {
            auto run = [&](auto tag) -> py::tuple {
                using T = decltype(tag);
                Csr<T> h = csr_from<T>(shape, indptr, indices, data);
                OptimizedHamiltonian<T> oh(&h, reorder);
                const Scale<> scale((double)energy.first, (double)energy.second);
                oh.optimize_for(Indices(src, dest), scale);
                py::array_t<int32_t> ip(oh.h2.indptr.size(), oh.h2.indptr.data());
                py::array_t<int32_t> ix(oh.h2.indices.size(), oh.h2.indices.data());
                py::array_t<T> dv(oh.h2.data.size(), oh.h2.data.data());
                py::array_t<int32_t> rm(oh.reorder_map.size(), oh.reorder_map.data());
                py::array_t<int32_t> bd(oh.slice_map.data.size(), oh.slice_map.data.data());
                py::array_t<int32_t> os(oh.optimized_idx.src.size(), oh.optimized_idx.src.data());
                py::array_t<int32_t> od(oh.optimized_idx.dest.size(), oh.optimized_idx.dest.data());
                return py::make_tuple(ip, ix, dv, rm, bd, py::make_tuple(os, od),
                                      py::make_tuple(oh.slice_map.src_offset, oh.slice_map.dest_offset),
                                      py::make_tuple(scale.a, scale.b));
            };
            auto dt = py::dtype::from_args(data.attr("dtype"));
            if (dt.kind() == 'f' && dt.itemsize() == 4) return run(float{});
            if (dt.kind() == 'f') return run(double{});
            if (dt.itemsize() == 8) return run(std::complex<float>{});
            return run(cd{});
        },
        py::arg("shape"), py::arg("indptr"), py::arg("indices"), py::arg("data"), py::arg("energy_range"),
        py::arg("reorder") = true, py::arg("src") = std::vector<idx_t>{0}, py::arg("dest") = std::vector<idx_t>{0});

    // Reference random starters (std::mt19937, default seed), `nvec` columns in original order,
    // returned vector-major and permuted by `reorder_map` (empty: identity); drawn `group`
    // columns per call of the starter (0: all at once).
    m.def(
        "random_starters",
        [](idx_t n, idx_t nvec, std::string dtype, py::array_t<int32_t, py::array::c_style | py::array::forcecast> reorder_map,
           idx_t group) -> py::array {
            if (reorder_map.size() != 0 && reorder_map.size() != n)
                throw std::invalid_argument("random_starters: reorder_map must be empty or of length n");
            auto run = [&](auto tag) -> py::array {
                using T = decltype(tag);
                Csr<T> dummy;
                dummy.rows = n;
                OptimizedHamiltonian<T> oh(&dummy, false);
                oh.reorder_map.assign(reorder_map.data(), reorder_map.data() + reorder_map.size());
                auto st = random_starter(oh);
                py::array_t<T> out({(py::ssize_t)nvec, (py::ssize_t)n});
                T* p = out.mutable_data();
                {
                    py::gil_scoped_release nogil;
                    const idx_t g = group > 0 ? group : std::max<idx_t>(nvec, 1);
                    for (idx_t v = 0; v < nvec; v += g) st.make_block(std::min(g, nvec - v), p + v * n);
                }
                return out;
            };
            if (dtype == "float32") return run(float{});
            if (dtype == "float64") return run(double{});
            if (dtype == "complex64") return run(std::complex<float>{});
            if (dtype == "complex128") return run(cd{});
            throw std::invalid_argument("unsupported dtype");
        },
        py::arg("n"), py::arg("nvec"), py::arg("dtype"),
        py::arg("reorder_map") = py::array_t<int32_t>(0), py::arg("group") = 0);

    // The reconstructions on their own (the KPM methods call them after the moments): KuboBastin
    // of damped M x M moments, SpectralDensity and GreensFunction of damped moments; the sums run
    // on `device`.
    auto real_or_complex = [](py::array mu, auto run) -> py::array {
        auto dt = py::dtype::from_args(mu.attr("dtype"));
        if (dt.kind() == 'f' && dt.itemsize() == 4) return run(float{});
        if (dt.kind() == 'f') return run(double{});
        if (dt.itemsize() == 8) return run(std::complex<float>{});
        return run(cd{});
    };
    m.def(
        "spectral_density",
        [real_or_complex](py::array mu, std::vector<double> energy, std::pair<double, double> ab, int device) {
            return real_or_complex(mu, [&](auto tag) -> py::array {
                using T = decltype(tag);
                auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(mu);
                if (!a || a.ndim() != 2) throw std::invalid_argument("mu must be cols x M");
                const idx_t cols = a.shape(0), M = a.shape(1);
                std::vector<T> v(a.data(), a.data() + cols * M);
                Scale<> sc;
                sc.a = ab.first;
                sc.b = ab.second;
                auto dev = Device::instance(device);
                std::vector<double> out;
                {
                    py::gil_scoped_release nogil;
                    out = spectral_density(dev->get(), v, M, cols, energy, sc);
                }
                py::array_t<double> r({(py::ssize_t)energy.size(), (py::ssize_t)cols});
                std::copy(out.begin(), out.end(), r.mutable_data());
                return r;
            });
        },
        py::arg("mu"), py::arg("energy"), py::arg("scale"), py::arg("device") = 0);
    m.def(
        "greens_function",
        [real_or_complex](py::array mu, std::vector<double> energy, std::pair<double, double> ab, int device) {
            return real_or_complex(mu, [&](auto tag) -> py::array {
                using T = decltype(tag);
                auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(mu);
                if (!a || a.ndim() != 1) throw std::invalid_argument("mu must be a vector");
                std::vector<T> v(a.data(), a.data() + a.shape(0));
                Scale<> sc;
                sc.a = ab.first;
                sc.b = ab.second;
                auto dev = Device::instance(device);
                std::vector<cd> out;
                {
                    py::gil_scoped_release nogil;
                    out = greens_function(dev->get(), v, energy, sc);
                }
                return py::array_t<cd>((py::ssize_t)out.size(), out.data());
            });
        },
        py::arg("mu"), py::arg("energy"), py::arg("scale"), py::arg("device") = 0);
    m.def(
        "kubo_bastin",
        [](py::array mu, std::vector<double> chemical_potential, std::vector<double> samples, double temperature,
           std::pair<double, double> ab, int device) -> py::array {
            auto run = [&](auto tag) -> py::array {
                using T = decltype(tag);
                auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(mu);
                if (!a || a.ndim() != 2 || a.shape(0) != a.shape(1)) throw std::invalid_argument("mu must be M x M");
                const idx_t M = a.shape(0);
                std::vector<T> v(a.data(), a.data() + M * M);
                Scale<> sc;
                sc.a = ab.first;
                sc.b = ab.second;
                auto dev = Device::instance(device);
                std::vector<cd> out;
                {
                    py::gil_scoped_release nogil;
                    out = kubo_bastin(dev->get(), v, M, chemical_potential, samples, temperature, sc);
                }
                return py::array_t<cd>((py::ssize_t)out.size(), out.data());
            };
            auto dt = py::dtype::from_args(mu.attr("dtype"));
            if (dt.kind() == 'f' && dt.itemsize() == 4) return run(float{});
            if (dt.kind() == 'f') return run(double{});
            if (dt.itemsize() == 8) return run(std::complex<float>{});
            return run(cd{});
        },
        py::arg("mu"), py::arg("chemical_potential"), py::arg("energy_samples"), py::arg("temperature"),
        py::arg("scale"), py::arg("device") = 0);
}
