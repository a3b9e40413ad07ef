// Host C++ state vector over the gfx950 engine's C-ABI (include/pl_amd.h).
//
// Mirrors the C++ surface Lightning-Qubit's measurements, observables and
// adjoint code consume (core/src/simulators/base/StateVectorBase.hpp:43-195,
// simulators/lightning_qubit/StateVectorLQubit.hpp:57-366): PrecisionT /
// ComplexT, getNumQubits, getLength, applyOperation(s), applyMatrix,
// applyGenerator, updateData, a copy constructor and construction from
// (pointer, length).  The amplitudes live in HBM; getDataVector() copies them
// back to the host.
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "pl_amd.h"

namespace Pennylane::LightningAMD {

/* Lightning-Qubit raises LightningException (PL_ABORT), std::invalid_argument
 * and std::out_of_range (core/src/utils/Error.hpp:22-128,
 * DynamicDispatcher.hpp:223-226,323-325); the C-ABI status codes map back to
 * the same C++ types. */
struct LightningException : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc) {
    if (rc == PL_OK) {
        return;
    }
    const std::string msg = pl_last_error();
    if (rc == PL_ERR_RANGE) {
        throw std::out_of_range(msg);
    }
    if (rc == PL_ERR_INVALID) {
        throw std::invalid_argument(msg);
    }
    throw LightningException(msg);
}

/* Gate names and ids: core/src/gates/Constant.hpp:46-96 / GateOperation.hpp:27-70 */
inline const std::vector<std::string> &gateNames() {
    static const std::vector<std::string> names = {
        "Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "PhaseShift",
        "RX", "RY", "RZ", "Rot", "CNOT", "CY", "CZ", "SWAP", "IsingXX", "IsingXY",
        "IsingYY", "IsingZZ", "ControlledPhaseShift", "CRX", "CRY", "CRZ", "CRot",
        "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus",
        "Toffoli", "CSWAP", "DoubleExcitation", "DoubleExcitationMinus",
        "DoubleExcitationPlus", "MultiRZ"};
    return names;
}
inline int gateId(const std::string &name) {
    const auto &n = gateNames();
    for (size_t i = 0; i < n.size(); i++) {
        if (n[i] == name) {
            return static_cast<int>(i);
        }
    }
    if (name == "QubitStateVector" || name == "BasisState") {
        return -1;
    }
    throw std::out_of_range("unordered_map::at: unknown operation " + name);
}

template <class PrecisionT_> class StateVectorAMD {
  public:
    using PrecisionT = PrecisionT_;
    using ComplexT = std::complex<PrecisionT>;
    static_assert(std::is_same_v<PrecisionT, float> || std::is_same_v<PrecisionT, double>);
    static constexpr int kDtype = std::is_same_v<PrecisionT, double> ? PL_C128 : PL_C64;

    /* |0...0> (StateVectorLQubitManaged.hpp:72-79) */
    explicit StateVectorAMD(size_t num_qubits) : num_qubits_(num_qubits) {
        check(pl_sv_create(static_cast<int>(num_qubits), kDtype, &h_));
    }
    /* copy of caller data (StateVectorLQubitRaw.hpp:78-87 wraps; here the
     * amplitudes are uploaded) */
    StateVectorAMD(const ComplexT *data, size_t length) : num_qubits_(log2_exact(length)) {
        check(pl_sv_create(static_cast<int>(num_qubits_), kDtype, &h_));
        updateData(data, length);
    }
    StateVectorAMD(const StateVectorAMD &other) : num_qubits_(other.num_qubits_) {
        check(pl_sv_create(static_cast<int>(num_qubits_), kDtype, &h_));
        check(pl_sv_copy(h_, other.h_));
    }
    StateVectorAMD &operator=(const StateVectorAMD &other) {
        if (this != &other) {
            if (other.num_qubits_ != num_qubits_) {
                throw std::invalid_argument("state vectors differ in size");
            }
            check(pl_sv_copy(h_, other.h_));
        }
        return *this;
    }
    StateVectorAMD(StateVectorAMD &&o) noexcept : num_qubits_(o.num_qubits_), h_(o.h_) {
        o.h_ = nullptr;
    }
    ~StateVectorAMD() {
        if (h_) {
            pl_sv_destroy(h_);
        }
    }

    [[nodiscard]] size_t getNumQubits() const { return num_qubits_; }
    [[nodiscard]] size_t getLength() const { return size_t{1} << num_qubits_; }
    [[nodiscard]] pl_sv_t handle() const { return h_; }
    /* queue gates and apply them as tile passes when the state is next used
     * (pl_sv_set_deferred); off applies the queue */
    void setDeferred(bool on) { check(pl_sv_set_deferred(h_, on ? 1 : 0)); }

    void applyOperation(const std::string &opName, const std::vector<size_t> &wires,
                        bool inverse = false, const std::vector<PrecisionT> &params = {}) {
        const int id = gateId(opName);
        if (id < 0) {
            throw std::out_of_range("state preparation is not a gate: " + opName);
        }
        const auto w = to_int(wires);
        const std::vector<double> p(params.begin(), params.end());
        check(pl_sv_apply_gate(h_, id, w.data(), static_cast<int>(w.size()), inverse ? 1 : 0,
                               p.data(), static_cast<int>(p.size())));
    }

    /* StateVectorBase::applyOperations (StateVectorBase.hpp:84-120) */
    void applyOperations(const std::vector<std::string> &ops,
                         const std::vector<std::vector<size_t>> &ops_wires,
                         const std::vector<bool> &ops_adjoint,
                         const std::vector<std::vector<PrecisionT>> &ops_params) {
        if (ops.size() != ops_wires.size() || ops.size() != ops_adjoint.size() ||
            ops.size() != ops_params.size()) {
            throw LightningException("Invalid arguments: number of operations, wires, inverses, "
                                     "and parameters must all be equal");
        }
        for (size_t i = 0; i < ops.size(); i++) {
            applyOperation(ops[i], ops_wires[i], ops_adjoint[i], ops_params[i]);
        }
    }
    void applyOperations(const std::vector<std::string> &ops,
                         const std::vector<std::vector<size_t>> &ops_wires,
                         const std::vector<bool> &ops_adjoint) {
        if (ops.size() != ops_wires.size() || ops.size() != ops_adjoint.size()) {
            throw LightningException("Invalid arguments: number of operations, wires, and "
                                     "inverses must all be equal");
        }
        for (size_t i = 0; i < ops.size(); i++) {
            applyOperation(ops[i], ops_wires[i], ops_adjoint[i], {});
        }
    }

    void applyMatrix(const ComplexT *matrix, const std::vector<size_t> &wires,
                     bool inverse = false) {
        if (wires.empty()) {
            throw LightningException("Number of wires must be larger than 0");
        }
        const size_t dim = size_t{1} << (2 * wires.size());
        std::vector<double> m(2 * dim);
        for (size_t i = 0; i < dim; i++) {
            m[2 * i] = matrix[i].real();
            m[2 * i + 1] = matrix[i].imag();
        }
        const auto w = to_int(wires);
        check(pl_sv_apply_matrix(h_, m.data(), w.data(), static_cast<int>(w.size()),
                                 inverse ? 1 : 0));
    }
    void applyMatrix(const std::vector<ComplexT> &matrix, const std::vector<size_t> &wires,
                     bool inverse = false) {
        if (matrix.size() != (size_t{1} << (2 * wires.size()))) {
            throw LightningException("The size of matrix does not match with the given "
                                     "number of wires");
        }
        applyMatrix(matrix.data(), wires, inverse);
    }

    [[nodiscard]] PrecisionT applyGenerator(const std::string &opName,
                                            const std::vector<size_t> &wires, bool adj = false) {
        const auto w = to_int(wires);
        double s = 0;
        check(pl_sv_apply_generator(h_, gateId(opName), w.data(), static_cast<int>(w.size()),
                                    adj ? 1 : 0, &s));
        return static_cast<PrecisionT>(s);
    }

    /* StateVectorLQubitManaged::updateData (StateVectorLQubitManaged.hpp:166-169) */
    void updateData(const ComplexT *data, size_t length) {
        if (length != getLength()) {
            throw LightningException("New data must be the same size as the old data.");
        }
        check(pl_sv_upload(h_, data, length * sizeof(ComplexT)));
    }
    void updateData(const std::vector<ComplexT> &data) { updateData(data.data(), data.size()); }
    void updateData(const StateVectorAMD &other) { check(pl_sv_copy(h_, other.h_)); }

    void getData(ComplexT *out) const { check(pl_sv_download(h_, out, getLength() * sizeof(ComplexT))); }
    [[nodiscard]] std::vector<ComplexT> getDataVector() const {
        std::vector<ComplexT> v(getLength());
        getData(v.data());
        return v;
    }
    void resetStateVector() { check(pl_sv_reset(h_)); }
    void setBasisState(size_t index) { check(pl_sv_set_basis_state(h_, index)); }

    static std::vector<int> to_int(const std::vector<size_t> &w) {
        return std::vector<int>(w.begin(), w.end());
    }

  private:
    static size_t log2_exact(size_t length) {
        size_t n = 0;
        while ((size_t{1} << n) < length) {
            n++;
        }
        if ((size_t{1} << n) != length || length < 2) {
            throw std::invalid_argument("state length must be a power of two >= 2");
        }
        return n;
    }
    size_t num_qubits_;
    pl_sv_t h_ = nullptr;
};

} // namespace Pennylane::LightningAMD
