// pybind11 module `lightning_qubit_ops`: the Python surface of Lightning-Qubit's
// native module (core/src/bindings/Bindings.cpp:23-57, Bindings.hpp:100-623,
// BindingsBase.hpp:39-83, simulators/lightning_qubit/bindings/LQubitBindings.hpp:45-264)
// with the gfx950 engine behind it.
//
// Residency: StateVectorC64/C128(ndarray) uploads the caller's buffer once;
// gates run on the device; the buffer is written back when the wrapper is
// destroyed or sync() is called, so the reference device's pattern
//   sim = StateVectorC128(state_vector); sim.RX(...); ...; return state_vector
// sees the updated amplitudes once `sim` goes out of scope.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <memory>
#include <span>
#include <string>
#include <vector>

#include "AdjointJacobianAMD.hpp"
#include "MeasurementsAMD.hpp"
#include "ObservablesAMD.hpp"
#include "StateVectorAMD.hpp"

namespace py = pybind11;
using namespace Pennylane::LightningAMD;
using namespace pybind11::literals;

/* core/src/utils/CPUMemoryModel.hpp:32-78 */
enum class CPUMemoryModel : uint8_t { Unaligned, Aligned256, Aligned512 };

template <class P> struct PyStateVector {
    using SV = StateVectorAMD<P>;
    py::object arr; // the caller's numpy buffer (or None for a device-only state)
    std::unique_ptr<SV> sv;
    bool dirty = false;

    explicit PyStateVector(const py::array_t<std::complex<P>> &a) : arr(a) {
        const py::buffer_info info = a.request();
        if (info.ndim != 1) {
            throw std::invalid_argument("NumPy array must be a 1-dimensional array");
        }
        if (info.itemsize != sizeof(std::complex<P>)) {
            throw std::invalid_argument("NumPy array must be of type np.complex64 or np.complex128");
        }
        sv = std::make_unique<SV>(static_cast<const std::complex<P> *>(info.ptr),
                                  static_cast<size_t>(info.shape[0]));
    }
    explicit PyStateVector(size_t num_qubits) : arr(py::none()), sv(std::make_unique<SV>(num_qubits)) {}
    ~PyStateVector() {
        try {
            sync();
        } catch (...) {
        }
    }
    void sync() {
        if (dirty && !arr.is_none()) {
            auto a = py::reinterpret_borrow<py::array_t<std::complex<P>>>(arr);
            sv->getData(a.mutable_data());
        }
        dirty = false;
    }
    py::array_t<std::complex<P>> state() const {
        py::array_t<std::complex<P>> out(sv->getLength());
        sv->getData(out.mutable_data());
        return out;
    }
};

template <class P> void registerPrecision(py::module_ &m, py::module_ &obs_m, py::module_ &alg_m) {
    using PySV = PyStateVector<P>;
    using SV = StateVectorAMD<P>;
    using ComplexT = std::complex<P>;
    using np_arr_c = py::array_t<ComplexT, py::array::c_style | py::array::forcecast>;
    using np_arr_r = py::array_t<P, py::array::c_style | py::array::forcecast>;
    using np_arr_idx = py::array_t<size_t, py::array::c_style | py::array::forcecast>;
    const std::string bits = std::to_string(sizeof(ComplexT) * 8);

    /* ---- StateVectorC* (Bindings.hpp:569-614, BindingsBase.hpp:39-83) ---- */
    auto cls = py::class_<PySV>(m, ("StateVectorC" + bits).c_str(), py::module_local())
                   .def(py::init<const py::array_t<ComplexT> &>())
                   .def_static("from_qubits", [](size_t n) { return std::make_unique<PySV>(n); },
                               "Device-resident |0...0> state with no host buffer.")
                   .def_property_readonly("size", [](const PySV &s) { return s.sv->getLength(); })
                   .def("num_qubits", [](const PySV &s) { return s.sv->getNumQubits(); })
                   .def("sync", &PySV::sync, "Write the device amplitudes back to the numpy buffer.")
                   .def("state", &PySV::state, "Copy of the device amplitudes.")
                   .def("resetStateVector", [](PySV &s) { s.sv->resetStateVector(); s.dirty = true; })
                   .def("setBasisState", [](PySV &s, size_t i) { s.sv->setBasisState(i); s.dirty = true; })
                   .def(
                       "applyMatrix",
                       [](PySV &s, const np_arr_c &matrix, const std::vector<size_t> &wires,
                          bool inverse) {
                           s.sv->applyMatrix(matrix.data(), wires, inverse);
                           s.dirty = true;
                       },
                       "Apply a given matrix to wires.", "matrix"_a, "wires"_a, "inverse"_a = false)
                   .def("kernel_map", [](const PySV &) {
                       py::dict d;
                       for (const auto &g : gateNames()) {
                           d[py::str(g)] = "gfx950";
                       }
                       return d;
                   });
    for (const auto &gname : gateNames()) {
        const std::string doc = "Apply the " + gname + " gate.";
        cls.def(
            gname.c_str(),
            [gname](PySV &s, const std::vector<size_t> &wires, bool inverse,
                    const std::vector<P> &params) {
                s.sv->applyOperation(gname, wires, inverse, params);
                s.dirty = true;
            },
            doc.c_str());
    }

    /* ---- observables (Bindings.hpp:276-395) ---- */
    using Obs = Observable<SV>;
    using ObsPtr = std::shared_ptr<Obs>;
    py::class_<Obs, ObsPtr>(obs_m, ("ObservableC" + bits).c_str(), py::module_local());
    auto eq = [](auto *tag) {
        using C = std::remove_pointer_t<decltype(tag)>;
        return [](const C &self, py::handle other) -> bool {
            if (!py::isinstance<C>(other)) {
                return false;
            }
            return self == other.cast<const C &>();
        };
    };
    py::class_<NamedObs<SV>, std::shared_ptr<NamedObs<SV>>, Obs>(obs_m, ("NamedObsC" + bits).c_str(),
                                                                 py::module_local())
        .def(py::init([](const std::string &name, const std::vector<size_t> &wires) {
            return NamedObs<SV>(name, wires);
        }))
        .def("__repr__", &NamedObs<SV>::getObsName)
        .def("get_wires", &NamedObs<SV>::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<NamedObs<SV> *>(nullptr)), "Compare two observables");
    py::class_<HermitianObs<SV>, std::shared_ptr<HermitianObs<SV>>, Obs>(
        obs_m, ("HermitianObsC" + bits).c_str(), py::module_local())
        .def(py::init([](const np_arr_c &matrix, const std::vector<size_t> &wires) {
            const ComplexT *p = matrix.data();
            return HermitianObs<SV>(std::vector<ComplexT>(p, p + matrix.size()), wires);
        }))
        .def("__repr__", &HermitianObs<SV>::getObsName)
        .def("get_wires", &HermitianObs<SV>::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<HermitianObs<SV> *>(nullptr)), "Compare two observables");
    py::class_<TensorProdObs<SV>, std::shared_ptr<TensorProdObs<SV>>, Obs>(
        obs_m, ("TensorProdObsC" + bits).c_str(), py::module_local())
        .def(py::init([](const std::vector<ObsPtr> &obs) { return TensorProdObs<SV>(obs); }))
        .def("__repr__", &TensorProdObs<SV>::getObsName)
        .def("get_wires", &TensorProdObs<SV>::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<TensorProdObs<SV> *>(nullptr)), "Compare two observables");
    py::class_<Hamiltonian<SV>, std::shared_ptr<Hamiltonian<SV>>, Obs>(
        obs_m, ("HamiltonianC" + bits).c_str(), py::module_local())
        .def(py::init([](const np_arr_r &coeffs, const std::vector<ObsPtr> &obs) {
            const P *p = coeffs.data();
            return Hamiltonian<SV>(std::vector<P>(p, p + coeffs.size()), obs);
        }))
        .def("__repr__", &Hamiltonian<SV>::getObsName)
        .def("get_wires", &Hamiltonian<SV>::getWires, "Get wires of observables")
        .def("__eq__", eq(static_cast<Hamiltonian<SV> *>(nullptr)), "Compare two observables");

    /* ---- MeasurementsC* (Bindings.hpp:405-451, LQubitBindings.hpp:107-183) ---- */
    using Meas = Measurements<SV>;
    struct PyMeas {
        PySV *owner;
        Meas m;
        explicit PyMeas(PySV &s) : owner(&s), m(*s.sv) {}
    };
    py::class_<PyMeas>(m, ("MeasurementsC" + bits).c_str(), py::module_local())
        .def(py::init<PySV &>(), py::keep_alive<1, 2>())
        .def("probs", [](PyMeas &M, const std::vector<size_t> &wires) {
            return py::array_t<P>(py::cast(M.m.probs(wires)));
        })
        .def("probs", [](PyMeas &M) { return py::array_t<P>(py::cast(M.m.probs())); })
        .def("expval", [](PyMeas &M, const ObsPtr &ob) { return M.m.expval(*ob); },
             "Expected value of an observable object.")
        .def("var", [](PyMeas &M, const ObsPtr &ob) { return M.m.var(*ob); },
             "Variance of an observable object.")
        .def("expval", [](PyMeas &M, const std::string &name, const std::vector<size_t> &wires) {
            return M.m.expval(name, wires);
        }, "Expected value of an operation by name.")
        .def("var", [](PyMeas &M, const std::string &name, const std::vector<size_t> &wires) {
            return M.m.var(name, wires);
        }, "Variance of an operation by name.")
        .def("generate_samples", [](PyMeas &M, size_t num_wires, size_t num_shots) {
            if (num_wires != M.owner->sv->getNumQubits()) {
                throw std::invalid_argument("num_wires must equal the number of qubits");
            }
            auto result = M.m.generate_samples(num_shots);
            py::array_t<size_t> out({num_shots, num_wires});
            std::copy(result.begin(), result.end(), out.mutable_data());
            return out;
        })
        .def("expval", [](PyMeas &M, const np_arr_idx &row_map, const np_arr_idx &entries, const np_arr_c &values) {
            return M.m.expval(row_map.data(), static_cast<size_t>(row_map.size()), entries.data(), values.data(),
                              static_cast<size_t>(values.size()));
        }, "Expected value of a sparse Hamiltonian.")
        .def("var", [](PyMeas &M, const np_arr_idx &row_map, const np_arr_idx &entries, const np_arr_c &values) {
            return M.m.var(row_map.data(), static_cast<size_t>(row_map.size()), entries.data(), values.data(),
                           static_cast<size_t>(values.size()));
        }, "Variance of a sparse Hamiltonian.")
        .def("generate_mcmc_samples", [](PyMeas &M, size_t num_wires, const std::string &kernelname,
                                         size_t num_burnin, size_t num_shots) {
            if (num_wires != M.owner->sv->getNumQubits()) {
                throw std::invalid_argument("num_wires must equal the number of qubits");
            }
            auto result = M.m.generate_samples_metropolis(kernelname, num_burnin, num_shots);
            py::array_t<size_t> out({num_shots, num_wires});
            std::copy(result.begin(), result.end(), out.mutable_data());
            return out;
        })
        .def("sample_indices", [](PyMeas &M, const py::array_t<double, py::array::c_style | py::array::forcecast> &u) {
            std::vector<double> uv(u.data(), u.data() + u.size());
            return py::array_t<int64_t>(py::cast(M.m.sample_indices(uv)));
        }, "Inverse-CDF basis indices for the given uniforms.")
        .def("set_seed", [](PyMeas &M, uint64_t s) { M.m.set_seed(s); });

    /* ---- algorithms (Bindings.hpp:485-561) ---- */
    using Ops = OpsData<SV>;
    py::class_<Ops>(alg_m, ("OpsStructC" + bits).c_str(), py::module_local())
        .def(py::init<const std::vector<std::string> &, const std::vector<std::vector<P>> &,
                      const std::vector<std::vector<size_t>> &, const std::vector<bool> &,
                      const std::vector<std::vector<ComplexT>> &>())
        .def("__repr__", [](const Ops &ops) {
            std::string s = "Operations: [";
            for (size_t i = 0; i < ops.getSize(); i++) {
                s += "{'name': " + ops.getOpsName()[i] + ", 'params': " + vec_str(ops.getOpsParams()[i]) +
                     ", 'inv': " + (ops.getOpsInverses()[i] ? "1" : "0") + "}" +
                     (i + 1 < ops.getSize() ? "," : "");
            }
            return s + "]";
        });
    alg_m.def(("create_ops_listC" + bits).c_str(),
              [](const std::vector<std::string> &names, const std::vector<std::vector<P>> &params,
                 const std::vector<std::vector<size_t>> &wires, const std::vector<bool> &inverses,
                 const std::vector<np_arr_c> &matrices) {
                  std::vector<std::vector<ComplexT>> conv(matrices.size());
                  for (size_t i = 0; i < matrices.size(); i++) {
                      conv[i].assign(matrices[i].data(), matrices[i].data() + matrices[i].size());
                  }
                  return Ops(names, params, wires, inverses, conv);
              },
              "Create a list of operations from data.");
    py::class_<AdjointJacobian<SV>>(alg_m, ("AdjointJacobianC" + bits).c_str(), py::module_local())
        .def(py::init<>())
        .def("__call__",
             [](AdjointJacobian<SV> &adj, const PySV &s, const std::vector<ObsPtr> &obs,
                const Ops &ops, const std::vector<size_t> &tp) {
                 std::vector<P> jac(obs.size() * tp.size(), P{0});
                 const JacobianData<SV> jd{ops.getTotalNumParams(), *s.sv, obs, ops, tp};
                 adj.adjointJacobian(std::span{jac}, jd);
                 return py::array_t<P>(py::cast(jac));
             },
             "Adjoint Jacobian method.");
    /* VectorJacobianProductC* (LQubitBindings.hpp:185-246) */
    py::class_<VectorJacobianProduct<SV>>(alg_m, ("VectorJacobianProductC" + bits).c_str(), py::module_local())
        .def(py::init<>())
        .def("__call__",
             [](VectorJacobianProduct<SV> &vjp, const PySV &s, const Ops &ops, const np_arr_c &dy,
                const std::vector<size_t> &tp) {
                 std::vector<ComplexT> out(tp.size(), ComplexT{});
                 const JacobianData<SV> jd{ops.getTotalNumParams(), *s.sv, {}, ops, tp};
                 vjp(std::span{out}, jd, std::span<const ComplexT>{dy.data(), static_cast<size_t>(dy.size())});
                 return py::array_t<ComplexT>(py::cast(out));
             },
             "Vector Jacobian Product method.");
}

PYBIND11_MODULE(lightning_qubit_ops, m) {
    m.doc() = "Lightning-Qubit compatible bindings over the MI355X (gfx950) engine";
    py::module_ obs_m = m.def_submodule("observables", "Submodule for observables classes.");
    py::module_ alg_m = m.def_submodule("algorithms", "Submodule for the algorithms functionality.");
    registerPrecision<float>(m, obs_m, alg_m);
    registerPrecision<double>(m, obs_m, alg_m);

    m.def("backend_info", [] { return py::dict("NAME"_a = "lightning.qubit", "DEVICE"_a = "gfx950"); },
          "Backend-specific information.");
    m.def("compile_info", [] {
        return py::dict("cpu.arch"_a = "x86_64", "compiler.name"_a = "GCC",
                        "compiler.version"_a = __VERSION__, "AVX2"_a = false, "AVX512F"_a = false,
                        "offload.arch"_a = "gfx950");
    }, "Compiled binary information.");
    m.def("runtime_info", [] {
        char buf[512];
        check(pl_device_info(buf, sizeof(buf)));
        return py::dict("AVX"_a = false, "AVX2"_a = false, "AVX512F"_a = false, "device"_a = std::string(buf));
    }, "Runtime information.");
    /* memory-alignment helpers the device class calls (Bindings.hpp:184-198,
     * lightning_qubit.py:233-239).  Host buffers are only staging for HBM
     * here, so no alignment is required: best_alignment() is Unaligned and
     * the device never reallocates a caller's array. */
    py::enum_<CPUMemoryModel>(m, "CPUMemoryModel", py::module_local())
        .value("Unaligned", CPUMemoryModel::Unaligned)
        .value("Aligned256", CPUMemoryModel::Aligned256)
        .value("Aligned512", CPUMemoryModel::Aligned512);
    m.def("best_alignment", [] { return CPUMemoryModel::Unaligned; },
          "Best memory alignment for the simulator.");
    m.def("get_alignment", [](const py::array &a) {
        const auto p = reinterpret_cast<uintptr_t>(a.request().ptr);
        if (p % 64 == 0) {
            return CPUMemoryModel::Aligned512;
        }
        return p % 32 == 0 ? CPUMemoryModel::Aligned256 : CPUMemoryModel::Unaligned;
    }, "Get alignment of an underlying data for a numpy array.");
    m.def("allocate_aligned_array", [](size_t size, const py::dtype &dt) {
        py::module_ np = py::module_::import("numpy");
        const size_t item = static_cast<size_t>(dt.itemsize());
        py::array raw = np.attr("zeros")(size * item + 64, "dtype"_a = "uint8");
        const auto p = reinterpret_cast<uintptr_t>(raw.request().ptr);
        const size_t off = (64 - p % 64) % 64;
        py::object sl = raw[py::slice(static_cast<py::ssize_t>(off),
                                      static_cast<py::ssize_t>(off + size * item), 1)];
        return sl.attr("view")(dt);
    }, "Get numpy array whose underlying data is aligned.");
}
