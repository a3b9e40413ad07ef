// Observable classes over the engine (core/src/observables/Observables.hpp:35-417,
// simulators/lightning_qubit/observables/ObservablesLQubit.hpp:48-355).
//
// Same names, constructor arguments, getObsName / getWires / operator== and
// error behaviour as Lightning-Qubit.  Each observable also serialises itself
// to the observable program of include/pl_amd.h, which is how measurements
// and the adjoint Jacobian hand it to the device.
#pragma once

#include <algorithm>
#include <complex>
#include <memory>
#include <sstream>
#include <string>
#include <typeinfo>
#include <unordered_set>
#include <vector>

#include "StateVectorAMD.hpp"

namespace Pennylane::LightningAMD {

struct ObsProgram {
    std::vector<int> prog;
    std::vector<double> coeffs;
    std::vector<double> mats;
    void begin(size_t nobs) {
        prog.assign(1, static_cast<int>(nobs));
        coeffs.clear();
        mats.clear();
    }
};

template <class T> std::string vec_str(const std::vector<T> &v) {
    std::ostringstream s;
    s << "[";
    for (size_t i = 0; i < v.size(); i++) {
        s << v[i] << (i + 1 < v.size() ? ", " : "");
    }
    s << "]";
    return s.str();
}

template <class StateVectorT> class Observable {
  public:
    using PrecisionT = typename StateVectorT::PrecisionT;
    virtual ~Observable() = default;
    virtual void applyInPlace(StateVectorT &sv) const = 0;
    [[nodiscard]] virtual std::string getObsName() const = 0;
    [[nodiscard]] virtual std::vector<size_t> getWires() const = 0;
    /* factor-level encoding: append the factors of one term */
    virtual void emitFactors(ObsProgram &p) const = 0;
    [[nodiscard]] virtual size_t numFactors() const = 0;
    /* observable-level encoding (is_ham, nterms, terms...) */
    virtual void emit(ObsProgram &p) const {
        p.prog.push_back(0);
        p.prog.push_back(1);
        p.prog.push_back(static_cast<int>(numFactors()));
        emitFactors(p);
    }
    [[nodiscard]] ObsProgram program() const {
        ObsProgram p;
        p.begin(1);
        emit(p);
        return p;
    }
    [[nodiscard]] bool operator==(const Observable &other) const {
        return typeid(*this) == typeid(other) && isEqual(other);
    }
    [[nodiscard]] bool operator!=(const Observable &other) const { return !(*this == other); }

  private:
    [[nodiscard]] virtual bool isEqual(const Observable &other) const = 0;
};

/* NamedObs: NamedObsBase, Observables.hpp:98-150 */
template <class StateVectorT> class NamedObs final : public Observable<StateVectorT> {
  public:
    using PrecisionT = typename StateVectorT::PrecisionT;
    NamedObs(std::string name, std::vector<size_t> wires, std::vector<PrecisionT> params = {})
        : name_(std::move(name)), wires_(std::move(wires)), params_(std::move(params)) {
        gateId(name_); // unknown names raise std::out_of_range, as the dispatcher would
    }
    void applyInPlace(StateVectorT &sv) const override {
        sv.applyOperation(name_, wires_, false, params_);
    }
    [[nodiscard]] std::string getObsName() const override { return name_ + vec_str(wires_); }
    [[nodiscard]] std::vector<size_t> getWires() const override { return wires_; }
    void emitFactors(ObsProgram &p) const override {
        p.prog.push_back(0);
        p.prog.push_back(gateId(name_));
        p.prog.push_back(static_cast<int>(wires_.size()));
        for (auto w : wires_) {
            p.prog.push_back(static_cast<int>(w));
        }
    }
    [[nodiscard]] size_t numFactors() const override { return 1; }

  private:
    bool isEqual(const Observable<StateVectorT> &o) const override {
        const auto &x = static_cast<const NamedObs &>(o);
        return name_ == x.name_ && wires_ == x.wires_ && params_ == x.params_;
    }
    std::string name_;
    std::vector<size_t> wires_;
    std::vector<PrecisionT> params_;
};

/* HermitianObs: HermitianObsBase, Observables.hpp:152-202 */
template <class StateVectorT> class HermitianObs final : public Observable<StateVectorT> {
  public:
    using ComplexT = typename StateVectorT::ComplexT;
    using MatrixT = std::vector<ComplexT>;
    HermitianObs(MatrixT matrix, std::vector<size_t> wires)
        : matrix_(std::move(matrix)), wires_(std::move(wires)) {
        if (matrix_.size() != (size_t{1} << (2 * wires_.size()))) {
            throw LightningException("Assertion failed: matrix_.size() == Util::exp2(2 * wires_.size())");
        }
    }
    void applyInPlace(StateVectorT &sv) const override { sv.applyMatrix(matrix_, wires_); }
    [[nodiscard]] std::string getObsName() const override { return "Hermitian"; }
    [[nodiscard]] std::vector<size_t> getWires() const override { return wires_; }
    [[nodiscard]] const MatrixT &getMatrix() const { return matrix_; }
    void emitFactors(ObsProgram &p) const override {
        p.prog.push_back(1);
        p.prog.push_back(0);
        p.prog.push_back(static_cast<int>(wires_.size()));
        for (auto w : wires_) {
            p.prog.push_back(static_cast<int>(w));
        }
        for (const auto &z : matrix_) {
            p.mats.push_back(z.real());
            p.mats.push_back(z.imag());
        }
    }
    [[nodiscard]] size_t numFactors() const override { return 1; }

  private:
    bool isEqual(const Observable<StateVectorT> &o) const override {
        const auto &x = static_cast<const HermitianObs &>(o);
        return matrix_ == x.matrix_ && wires_ == x.wires_;
    }
    MatrixT matrix_;
    std::vector<size_t> wires_;
};

/* TensorProdObs: TensorProdObsBase, Observables.hpp:204-324 */
template <class StateVectorT> class TensorProdObs final : public Observable<StateVectorT> {
  public:
    using ObsPtr = std::shared_ptr<Observable<StateVectorT>>;
    explicit TensorProdObs(std::vector<ObsPtr> obs) : obs_(std::move(obs)) {
        std::unordered_set<size_t> seen;
        for (const auto &ob : obs_) {
            if (ob->numFactors() == 0) {
                throw LightningException("a Hamiltonian cannot be a tensor-product factor");
            }
            for (auto w : ob->getWires()) {
                if (seen.count(w)) {
                    throw LightningException("All wires in observables must be disjoint.");
                }
                seen.insert(w);
            }
        }
        all_wires_.assign(seen.begin(), seen.end());
        std::sort(all_wires_.begin(), all_wires_.end());
    }
    static std::shared_ptr<TensorProdObs> create(std::vector<ObsPtr> obs) {
        return std::make_shared<TensorProdObs>(std::move(obs));
    }
    void applyInPlace(StateVectorT &sv) const override {
        for (const auto &ob : obs_) {
            ob->applyInPlace(sv);
        }
    }
    [[nodiscard]] std::string getObsName() const override {
        std::string s;
        for (size_t i = 0; i < obs_.size(); i++) {
            s += obs_[i]->getObsName() + (i + 1 < obs_.size() ? " @ " : "");
        }
        return s;
    }
    [[nodiscard]] std::vector<size_t> getWires() const override { return all_wires_; }
    [[nodiscard]] size_t getSize() const { return obs_.size(); }
    void emitFactors(ObsProgram &p) const override {
        for (const auto &ob : obs_) {
            ob->emitFactors(p);
        }
    }
    [[nodiscard]] size_t numFactors() const override {
        size_t k = 0;
        for (const auto &ob : obs_) {
            k += ob->numFactors();
        }
        return k;
    }

  private:
    bool isEqual(const Observable<StateVectorT> &o) const override {
        const auto &x = static_cast<const TensorProdObs &>(o);
        if (obs_.size() != x.obs_.size()) {
            return false;
        }
        for (size_t i = 0; i < obs_.size(); i++) {
            if (*obs_[i] != *x.obs_[i]) {
                return false;
            }
        }
        return true;
    }
    std::vector<ObsPtr> obs_;
    std::vector<size_t> all_wires_;
};

/* Hamiltonian: HamiltonianBase (Observables.hpp:326-417) with the device
 * applyInPlace of res = sum_t c_t O_t psi (ObservablesLQubit.hpp:156-190). */
template <class StateVectorT> class Hamiltonian final : public Observable<StateVectorT> {
  public:
    using PrecisionT = typename StateVectorT::PrecisionT;
    using ObsPtr = std::shared_ptr<Observable<StateVectorT>>;
    Hamiltonian(std::vector<PrecisionT> coeffs, std::vector<ObsPtr> obs)
        : coeffs_(std::move(coeffs)), obs_(std::move(obs)) {
        if (coeffs_.size() != obs_.size()) {
            throw LightningException("Assertion failed: coeffs_.size() == obs_.size()");
        }
        for (const auto &o : obs_) {
            if (dynamic_cast<const Hamiltonian *>(o.get())) {
                throw LightningException("nested Hamiltonian terms are not supported");
            }
        }
    }
    static std::shared_ptr<Hamiltonian> create(std::vector<PrecisionT> coeffs, std::vector<ObsPtr> obs) {
        return std::make_shared<Hamiltonian>(std::move(coeffs), std::move(obs));
    }
    void applyInPlace(StateVectorT &sv) const override {
        const ObsProgram p = this->program();
        check(pl_sv_apply_observable(sv.handle(), p.prog.data(), p.coeffs.data(), p.mats.data()));
    }
    [[nodiscard]] std::vector<size_t> getWires() const override {
        std::unordered_set<size_t> w;
        for (const auto &o : obs_) {
            for (auto x : o->getWires()) {
                w.insert(x);
            }
        }
        std::vector<size_t> v(w.begin(), w.end());
        std::sort(v.begin(), v.end());
        return v;
    }
    [[nodiscard]] std::string getObsName() const override {
        std::ostringstream ss;
        ss << "Hamiltonian: { 'coeffs' : " << vec_str(coeffs_) << ", 'observables' : [";
        for (size_t t = 0; t < coeffs_.size(); t++) {
            ss << obs_[t]->getObsName() << (t + 1 < coeffs_.size() ? ", " : "");
        }
        ss << "]}";
        return ss.str();
    }
    void emitFactors(ObsProgram &) const override {
        throw LightningException("a Hamiltonian cannot be a tensor-product factor");
    }
    [[nodiscard]] size_t numFactors() const override { return 0; }
    void emit(ObsProgram &p) const override {
        p.prog.push_back(1);
        p.prog.push_back(static_cast<int>(obs_.size()));
        for (size_t t = 0; t < obs_.size(); t++) {
            p.coeffs.push_back(static_cast<double>(coeffs_[t]));
            p.prog.push_back(static_cast<int>(obs_[t]->numFactors()));
            obs_[t]->emitFactors(p);
        }
    }

  private:
    bool isEqual(const Observable<StateVectorT> &o) const override {
        const auto &x = static_cast<const Hamiltonian &>(o);
        if (coeffs_ != x.coeffs_ || obs_.size() != x.obs_.size()) {
            return false;
        }
        for (size_t i = 0; i < obs_.size(); i++) {
            if (*obs_[i] != *x.obs_[i]) {
                return false;
            }
        }
        return true;
    }
    std::vector<PrecisionT> coeffs_;
    std::vector<ObsPtr> obs_;
};

} // namespace Pennylane::LightningAMD
