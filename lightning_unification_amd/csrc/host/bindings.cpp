// pybind11 module `lightning_qubit_ops`: the Python surface of Lightning-Qubit's
// native module (core/src/bindings/Bindings.cpp:23-57, Bindings.hpp:100-623,
// BindingsBase.hpp:39-83, simulators/lightning_qubit/bindings/LQubitBindings.hpp:45-264)
// with the gfx950 engine behind it.
//
// Residency (device mirror of the caller's buffer).  The reference's
// StateVectorC*(ndarray) is a view of the numpy buffer
// (StateVectorLQubitRaw.hpp:78-87) and the device wraps the same buffer again
// for every call (lightning_qubit.py:342-344, 430, 488, 528, 629).  Here the
// amplitudes live in HBM, in one Mirror per caller buffer, keyed by (data
// pointer, byte length) within a precision and shared by every wrapper of
// that buffer:
//   - the first wrap uploads the buffer; later wraps of an unchanged buffer
//     reuse the device copy and move no bytes;
//   - host writes into the buffer are detected exactly (HostWriteWatch.hpp);
//     the next call through any wrapper re-uploads the buffer first;
//   - gates change the device copy only; the buffer is written back when a
//     wrapper is destroyed (the device's `sim` at the end of
//     apply_lightning) or on sync();
//   - gates are deferred (pl_sv_set_deferred): the gates of apply_lightning,
//     one call each, queue up and run as tile passes when the state is next
//     used (a measurement, the adjoint, the write-back), so a circuit reads
//     and writes the state once per pass instead of once per gate;
//   - a call finding both the device copy changed by gates and the buffer
//     written on the host raises instead of choosing one of them;
//   - the mirror is freed when the buffer's owning array dies (a weak
//     reference with a callback), so a reused address never meets a stale
//     copy; a buffer whose owner takes no weak reference gets a mirror that
//     lives only as long as its wrappers.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <map>
#include <memory>
#include <span>
#include <string>
#include <vector>

#include "AdjointJacobianAMD.hpp"
#include "HostWriteWatch.hpp"
#include "MeasurementsAMD.hpp"
#include "ObservablesAMD.hpp"
#include "StateVectorAMD.hpp"

namespace py = pybind11;
using namespace Pennylane::LightningAMD;
using namespace pybind11::literals;

/* core/src/utils/CPUMemoryModel.hpp:32-78 */
enum class CPUMemoryModel : uint8_t { Unaligned, Aligned256, Aligned512 };

using hostwatch::WriteWatch;

template <class P> struct Mirror {
    std::unique_ptr<StateVectorAMD<P>> sv;
    void *host = nullptr; // caller's buffer (nullptr: device-only state)
    size_t nbytes = 0;
    WriteWatch watch;
    bool device_newer = false; // changed by gates since host and device agreed
    std::pair<uintptr_t, size_t> key{0, 0};
    bool registered = false;

    // before a call reads or changes the device copy
    void prepare() {
        if (!host || !watch.written()) {
            return;
        }
        if (device_newer) {
            throw LightningException("the state buffer was written on the host while the device copy "
                                     "holds newer amplitudes; call sync() before writing the buffer");
        }
        watch.release();
        sv->updateData(static_cast<const std::complex<P> *>(host), sv->getLength());
        watch.watch(host, nbytes);
    }
    void write_back() {
        if (!host || !device_newer) {
            return;
        }
        watch.release();
        sv->getData(static_cast<std::complex<P> *>(host));
        watch.watch(host, nbytes);
        device_newer = false;
    }
};

template <class P> struct MirrorRegistry {
    // (data pointer, bytes) -> mirror; owned here while the buffer's owner lives
    std::map<std::pair<uintptr_t, size_t>, std::shared_ptr<Mirror<P>>> live;
    std::map<std::pair<uintptr_t, size_t>, PyObject *> refs; // weak references to the owners
    static MirrorRegistry &get() {
        static auto *r = new MirrorRegistry(); // never destroyed after the interpreter
        return *r;
    }
    std::vector<PyObject *> dead; // weak references whose callback has run
    // runs inside the owner's weak-reference callback: the weak reference
    // itself is released later, outside its own callback
    void drop(const std::pair<uintptr_t, size_t> &k) {
        auto it = live.find(k);
        if (it != live.end()) {
            it->second->watch.release();
            it->second->registered = false;
            live.erase(it);
        }
        auto jt = refs.find(k);
        if (jt != refs.end()) {
            dead.push_back(jt->second);
            refs.erase(jt);
        }
    }
    void release_dead() {
        for (PyObject *o : dead) {
            Py_XDECREF(o);
        }
        dead.clear();
    }
    void clear() {
        while (!live.empty()) {
            drop(live.begin()->first);
        }
        release_dead();
    }
};

// the object that owns a numpy array's memory: follow .base while it is an array
inline py::object buffer_owner(const py::array &a) {
    py::object o = a;
    while (py::isinstance<py::array>(o)) {
        py::object b = o.attr("base");
        if (b.is_none()) {
            break;
        }
        o = b;
    }
    return o;
}

template <class P> struct PyStateVector {
    using SV = StateVectorAMD<P>;
    py::object arr; // the caller's numpy buffer (None for a device-only state)
    std::shared_ptr<Mirror<P>> mir;

    explicit PyStateVector(const py::array_t<std::complex<P>> &a) : arr(a) {
        const py::buffer_info info = a.request();
        if (info.ndim != 1) {
            throw std::invalid_argument("NumPy array must be a 1-dimensional array");
        }
        if (info.itemsize != sizeof(std::complex<P>)) {
            throw std::invalid_argument("NumPy array must be of type np.complex64 or np.complex128");
        }
        if (info.strides[0] != static_cast<py::ssize_t>(sizeof(std::complex<P>))) {
            throw std::invalid_argument("NumPy array must be contiguous");
        }
        const size_t len = static_cast<size_t>(info.shape[0]);
        const auto key = std::make_pair(reinterpret_cast<uintptr_t>(info.ptr), len * sizeof(std::complex<P>));
        auto &reg = MirrorRegistry<P>::get();
        reg.release_dead();
        auto it = reg.live.find(key);
        if (it != reg.live.end()) {
            mir = it->second;
            mir->prepare();
            return;
        }
        mir = std::make_shared<Mirror<P>>();
        mir->host = info.ptr;
        mir->nbytes = key.second;
        mir->key = key;
        mir->sv = std::make_unique<SV>(static_cast<const std::complex<P> *>(info.ptr), len);
        mir->sv->setDeferred(true); // a circuit's gates run as tile passes
        mir->watch.watch(info.ptr, key.second);
        // register under a weak reference to the buffer's owner
        py::object owner = buffer_owner(a);
        py::cpp_function on_death([key](py::handle) { MirrorRegistry<P>::get().drop(key); });
        PyObject *wr = PyWeakref_NewRef(owner.ptr(), on_death.ptr());
        if (wr) {
            reg.live[key] = mir;
            reg.refs[key] = wr;
            mir->registered = true;
        } else {
            PyErr_Clear(); // owner takes no weak reference: the mirror lives with its wrappers
        }
    }
    explicit PyStateVector(size_t num_qubits) : arr(py::none()), mir(std::make_shared<Mirror<P>>()) {
        mir->sv = std::make_unique<SV>(num_qubits);
        mir->sv->setDeferred(true);
    }
    ~PyStateVector() {
        try {
            sync();
        } catch (...) {
        }
    }
    SV &sv() const { return *mir->sv; }
    // before a call that reads or changes the device copy
    SV &use() const {
        mir->prepare();
        return *mir->sv;
    }
    // after a call that changed the device copy
    void touched() const { mir->device_newer = mir->host != nullptr; }
    void sync() const { mir->write_back(); }
    py::array_t<std::complex<P>> state() const {
        py::array_t<std::complex<P>> out(sv().getLength());
        use().getData(out.mutable_data());
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
                   .def_property_readonly("size", [](const PySV &s) { return s.sv().getLength(); })
                   .def("num_qubits", [](const PySV &s) { return s.sv().getNumQubits(); })
                   .def("sync", &PySV::sync, "Write the device amplitudes back to the numpy buffer.")
                   .def("state", &PySV::state, "Copy of the device amplitudes.")
                   .def("resetStateVector", [](PySV &s) { s.use().resetStateVector(); s.touched(); })
                   .def("setBasisState", [](PySV &s, size_t i) { s.use().setBasisState(i); s.touched(); })
                   .def(
                       "applyMatrix",
                       [](PySV &s, const np_arr_c &matrix, const std::vector<size_t> &wires,
                          bool inverse) {
                           s.use().applyMatrix(matrix.data(), wires, inverse);
                           s.touched();
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
                s.use().applyOperation(gname, wires, inverse, params);
                s.touched();
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
        explicit PyMeas(PySV &s) : owner(&s), m(s.sv()) {}
        Meas &use() {
            owner->use();
            return m;
        }
    };
    py::class_<PyMeas>(m, ("MeasurementsC" + bits).c_str(), py::module_local())
        .def(py::init<PySV &>(), py::keep_alive<1, 2>())
        .def("probs", [](PyMeas &M, const std::vector<size_t> &wires) {
            return py::array_t<P>(py::cast(M.use().probs(wires)));
        })
        .def("probs", [](PyMeas &M) { return py::array_t<P>(py::cast(M.use().probs())); })
        .def("expval", [](PyMeas &M, const ObsPtr &ob) { return M.use().expval(*ob); },
             "Expected value of an observable object.")
        .def("var", [](PyMeas &M, const ObsPtr &ob) { return M.use().var(*ob); },
             "Variance of an observable object.")
        .def("expval", [](PyMeas &M, const std::string &name, const std::vector<size_t> &wires) {
            return M.use().expval(name, wires);
        }, "Expected value of an operation by name.")
        .def("var", [](PyMeas &M, const std::string &name, const std::vector<size_t> &wires) {
            return M.use().var(name, wires);
        }, "Variance of an operation by name.")
        .def("generate_samples", [](PyMeas &M, size_t num_wires, size_t num_shots) {
            if (num_wires != M.owner->sv().getNumQubits()) {
                throw std::invalid_argument("num_wires must equal the number of qubits");
            }
            auto result = M.use().generate_samples(num_shots);
            py::array_t<size_t> out({num_shots, num_wires});
            std::copy(result.begin(), result.end(), out.mutable_data());
            return out;
        })
        .def("expval", [](PyMeas &M, const np_arr_idx &row_map, const np_arr_idx &entries, const np_arr_c &values) {
            return M.use().expval(row_map.data(), static_cast<size_t>(row_map.size()), entries.data(), values.data(),
                              static_cast<size_t>(values.size()));
        }, "Expected value of a sparse Hamiltonian.")
        .def("var", [](PyMeas &M, const np_arr_idx &row_map, const np_arr_idx &entries, const np_arr_c &values) {
            return M.use().var(row_map.data(), static_cast<size_t>(row_map.size()), entries.data(), values.data(),
                           static_cast<size_t>(values.size()));
        }, "Variance of a sparse Hamiltonian.")
        .def("generate_mcmc_samples", [](PyMeas &M, size_t num_wires, const std::string &kernelname,
                                         size_t num_burnin, size_t num_shots) {
            if (num_wires != M.owner->sv().getNumQubits()) {
                throw std::invalid_argument("num_wires must equal the number of qubits");
            }
            auto result = M.use().generate_samples_metropolis(kernelname, num_burnin, num_shots);
            py::array_t<size_t> out({num_shots, num_wires});
            std::copy(result.begin(), result.end(), out.mutable_data());
            return out;
        })
        .def("sample_indices", [](PyMeas &M, const py::array_t<double, py::array::c_style | py::array::forcecast> &u) {
            std::vector<double> uv(u.data(), u.data() + u.size());
            return py::array_t<int64_t>(py::cast(M.use().sample_indices(uv)));
        }, "Inverse-CDF basis indices for the given uniforms.")
        .def("set_seed", [](PyMeas &M, uint64_t s) { M.use().set_seed(s); });

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
                 const JacobianData<SV> jd{ops.getTotalNumParams(), s.use(), obs, ops, tp};
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
                 const JacobianData<SV> jd{ops.getTotalNumParams(), s.use(), {}, ops, tp};
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

    m.def("transfer_counters", [] {
        uint64_t h2d = 0, d2h = 0;
        check(pl_transfer_counters(&h2d, &d2h));
        return py::make_tuple(h2d, d2h);
    }, "Bytes moved host -> device and device -> host by state uploads and write-backs so far.");
    m.def("mirror_count", [] {
        return MirrorRegistry<float>::get().live.size() + MirrorRegistry<double>::get().live.size();
    }, "Number of caller buffers that currently have a device copy.");
    // device copies are released while the interpreter and the HIP runtime are
    // both still alive
    py::module_::import("atexit").attr("register")(py::cpp_function([] {
        MirrorRegistry<float>::get().clear();
        MirrorRegistry<double>::get().clear();
    }));
    /* The host write detection on its own, for tests that need no device. */
    py::class_<WriteWatch>(m, "_HostWriteWatch", py::module_local())
        .def(py::init<>())
        .def("watch", [](WriteWatch &w, const py::array &a) {
            const py::buffer_info info = a.request();
            w.watch(info.ptr, static_cast<size_t>(info.size * info.itemsize));
        })
        .def("written", &WriteWatch::written)
        .def("release", &WriteWatch::release);

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
