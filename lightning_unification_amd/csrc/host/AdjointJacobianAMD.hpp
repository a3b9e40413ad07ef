// Tape containers and the adjoint Jacobian over the engine
// (core/src/algorithms/JacobianData.hpp:40-335,
//  simulators/lightning_qubit/algorithms/AdjointJacobianLQubit.hpp:212-336).
#pragma once

#include <complex>
#include <memory>
#include <numeric>
#include <span>
#include <string>
#include <vector>

#include "ObservablesAMD.hpp"
#include "StateVectorAMD.hpp"

namespace Pennylane::LightningAMD {

template <class StateVectorT> class OpsData {
  public:
    using PrecisionT = typename StateVectorT::PrecisionT;
    using ComplexT = typename StateVectorT::ComplexT;

    OpsData(std::vector<std::string> ops_name, const std::vector<std::vector<PrecisionT>> &ops_params,
            std::vector<std::vector<size_t>> ops_wires, std::vector<bool> ops_inverses,
            std::vector<std::vector<ComplexT>> ops_matrices = {})
        : ops_name_(std::move(ops_name)), ops_params_(ops_params), ops_wires_(std::move(ops_wires)),
          ops_inverses_(std::move(ops_inverses)), ops_matrices_(std::move(ops_matrices)) {
        if (ops_matrices_.empty()) {
            ops_matrices_.resize(ops_name_.size());
        }
        for (const auto &p : ops_params_) {
            num_par_ops_ += p.empty() ? 0 : 1;
        }
        num_nonpar_ops_ = ops_params_.size() - num_par_ops_;
    }
    [[nodiscard]] size_t getSize() const { return ops_name_.size(); }
    [[nodiscard]] const std::vector<std::string> &getOpsName() const { return ops_name_; }
    [[nodiscard]] const std::vector<std::vector<PrecisionT>> &getOpsParams() const { return ops_params_; }
    [[nodiscard]] const std::vector<std::vector<size_t>> &getOpsWires() const { return ops_wires_; }
    [[nodiscard]] const std::vector<bool> &getOpsInverses() const { return ops_inverses_; }
    [[nodiscard]] const std::vector<std::vector<ComplexT>> &getOpsMatrices() const { return ops_matrices_; }
    [[nodiscard]] bool hasParams(size_t i) const { return !ops_params_[i].empty(); }
    [[nodiscard]] size_t getNumParOps() const { return num_par_ops_; }
    [[nodiscard]] size_t getNumNonParOps() const { return num_nonpar_ops_; }
    [[nodiscard]] size_t getTotalNumParams() const {
        size_t s = 0;
        for (const auto &p : ops_params_) {
            s += p.size();
        }
        return s;
    }

    /* tape encoding of include/pl_amd.h */
    struct Encoded {
        std::vector<int> ids, nwires, wires, nparams;
        std::vector<uint8_t> inverse;
        std::vector<double> params;
    };
    [[nodiscard]] Encoded encode() const {
        Encoded e;
        for (size_t i = 0; i < ops_name_.size(); i++) {
            e.ids.push_back(gateId(ops_name_[i]));
            e.nwires.push_back(static_cast<int>(ops_wires_[i].size()));
            for (auto w : ops_wires_[i]) {
                e.wires.push_back(static_cast<int>(w));
            }
            e.inverse.push_back(ops_inverses_[i] ? 1 : 0);
            e.nparams.push_back(static_cast<int>(ops_params_[i].size()));
            for (auto p : ops_params_[i]) {
                e.params.push_back(static_cast<double>(p));
            }
        }
        return e;
    }

  private:
    size_t num_par_ops_ = 0;
    size_t num_nonpar_ops_ = 0;
    std::vector<std::string> ops_name_;
    std::vector<std::vector<PrecisionT>> ops_params_;
    std::vector<std::vector<size_t>> ops_wires_;
    std::vector<bool> ops_inverses_;
    std::vector<std::vector<ComplexT>> ops_matrices_;
};

template <class StateVectorT> class JacobianData {
  public:
    using ObsPtr = std::shared_ptr<Observable<StateVectorT>>;
    JacobianData(size_t num_params, const StateVectorT &sv, std::vector<ObsPtr> obs,
                 OpsData<StateVectorT> ops, std::vector<size_t> trainP)
        : num_parameters_(num_params), sv_(sv), observables_(std::move(obs)),
          operations_(std::move(ops)), trainable_(std::move(trainP)) {}
    [[nodiscard]] size_t getNumParams() const { return num_parameters_; }
    [[nodiscard]] const StateVectorT &getStateVector() const { return sv_; }
    [[nodiscard]] size_t getSizeStateVec() const { return sv_.getLength(); }
    [[nodiscard]] const std::vector<ObsPtr> &getObservables() const { return observables_; }
    [[nodiscard]] size_t getNumObservables() const { return observables_.size(); }
    [[nodiscard]] const OpsData<StateVectorT> &getOperations() const { return operations_; }
    [[nodiscard]] const std::vector<size_t> &getTrainableParams() const { return trainable_; }
    [[nodiscard]] bool hasTrainableParams() const { return !trainable_.empty(); }

  private:
    size_t num_parameters_;
    const StateVectorT &sv_;
    std::vector<ObsPtr> observables_;
    OpsData<StateVectorT> operations_;
    std::vector<size_t> trainable_;
};

template <class StateVectorT> class AdjointJacobian {
  public:
    using PrecisionT = typename StateVectorT::PrecisionT;

    /* jac is observable-major, jac[k * |tp| + t] (AdjointJacobianLQubit.hpp:331-334).
     * apply_operations = true first applies the tape to a copy of the state,
     * as the reference does for a state prepared elsewhere. */
    void adjointJacobian(std::span<PrecisionT> jac, const JacobianData<StateVectorT> &jd,
                         bool apply_operations = false) {
        const auto &tp = jd.getTrainableParams();
        const auto &obs = jd.getObservables();
        if (!jd.hasTrainableParams()) {
            return;
        }
        if (jac.size() != tp.size() * obs.size()) {
            throw LightningException("The size of preallocated jacobian must be same as the number "
                                     "of trainable parameters times the number of observables "
                                     "provided.");
        }
        ObsProgram p;
        p.begin(obs.size());
        for (const auto &o : obs) {
            o->emit(p);
        }
        const auto e = jd.getOperations().encode();
        const std::vector<int> tpi(tp.begin(), tp.end());
        std::vector<double> out(jac.size());
        const StateVectorT *start = &jd.getStateVector();
        std::unique_ptr<StateVectorT> applied;
        if (apply_operations) {
            applied = std::make_unique<StateVectorT>(jd.getStateVector());
            check(pl_sv_apply_ops(applied->handle(), static_cast<int>(e.ids.size()), e.ids.data(),
                                  e.nwires.data(), e.wires.data(), e.inverse.data(),
                                  e.nparams.data(), e.params.data()));
            start = applied.get();
        }
        check(pl_adjoint_jacobian(start->handle(), static_cast<int>(e.ids.size()), e.ids.data(),
                                  e.nwires.data(), e.wires.data(), e.inverse.data(),
                                  e.nparams.data(), e.params.data(), p.prog.data(),
                                  p.coeffs.data(), p.mats.data(), static_cast<int>(tpi.size()),
                                  tpi.data(), out.data()));
        for (size_t i = 0; i < jac.size(); i++) {
            jac[i] = static_cast<PrecisionT>(out[i]);
        }
    }
};

/* VectorJacobianProduct (VectorJacobianProduct.hpp:42-157): w = dy^dagger J
 * for the state-vector Jacobian J_ij = d psi_i / d theta_j. */
template <class StateVectorT> class VectorJacobianProduct {
  public:
    using PrecisionT = typename StateVectorT::PrecisionT;
    using ComplexT = std::complex<PrecisionT>;

    void operator()(std::span<ComplexT> jac, const JacobianData<StateVectorT> &jd,
                    std::span<const ComplexT> dy, bool apply_operations = false) {
        const auto &tp = jd.getTrainableParams();
        if (dy.size() != jd.getStateVector().getLength()) {
            throw LightningException("Assertion failed: dy.size() == jd.getSizeStateVec()");
        }
        if (tp.empty()) {
            return;
        }
        if (jac.size() != tp.size()) {
            throw LightningException("The size of preallocated jacobian must be same as the number of "
                                     "trainable parameters.");
        }
        const auto e = jd.getOperations().encode();
        const std::vector<int> tpi(tp.begin(), tp.end());
        std::vector<double> d(2 * dy.size()), out(2 * tp.size());
        for (size_t i = 0; i < dy.size(); i++) {
            d[2 * i] = dy[i].real();
            d[2 * i + 1] = dy[i].imag();
        }
        const StateVectorT *start = &jd.getStateVector();
        std::unique_ptr<StateVectorT> applied;
        if (apply_operations) {
            applied = std::make_unique<StateVectorT>(jd.getStateVector());
            check(pl_sv_apply_ops(applied->handle(), static_cast<int>(e.ids.size()), e.ids.data(),
                                  e.nwires.data(), e.wires.data(), e.inverse.data(), e.nparams.data(),
                                  e.params.data()));
            start = applied.get();
        }
        check(pl_vjp(start->handle(), static_cast<int>(e.ids.size()), e.ids.data(), e.nwires.data(),
                     e.wires.data(), e.inverse.data(), e.nparams.data(), e.params.data(), d.data(),
                     static_cast<int>(tpi.size()), tpi.data(), out.data()));
        for (size_t t = 0; t < jac.size(); t++) {
            jac[t] = ComplexT(static_cast<PrecisionT>(out[2 * t]), static_cast<PrecisionT>(out[2 * t + 1]));
        }
    }
};

} // namespace Pennylane::LightningAMD
