// Measurements over the engine (core/src/measurements/MeasurementsBase.hpp:42-106,
// simulators/lightning_qubit/measurements/MeasurementsLQubit.hpp:60-641).
#pragma once

#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "ObservablesAMD.hpp"
#include "StateVectorAMD.hpp"

namespace Pennylane::LightningAMD {

template <class StateVectorT> class Measurements final {
  public:
    using PrecisionT = typename StateVectorT::PrecisionT;
    using ComplexT = typename StateVectorT::ComplexT;

    explicit Measurements(const StateVectorT &sv) : sv_(sv) {}

    /* probs(): MeasurementsLQubit.hpp:79-89 */
    std::vector<PrecisionT> probs() {
        std::vector<double> p(sv_.getLength());
        check(pl_probs(sv_.handle(), nullptr, -1, p.data()));
        return {p.begin(), p.end()};
    }
    /* probs(wires), axes ordered as the reference's sorted marginal followed
     * by transpose_state_tensor(.., sorting_indices(wires)):
     * MeasurementsLQubit.hpp:99-135 */
    std::vector<PrecisionT> probs(const std::vector<size_t> &wires,
                                  [[maybe_unused]] const std::vector<size_t> &device_wires = {}) {
        const auto w = StateVectorT::to_int(wires);
        std::vector<double> p(size_t{1} << wires.size());
        check(pl_probs(sv_.handle(), w.data(), static_cast<int>(w.size()), p.data()));
        return {p.begin(), p.end()};
    }

    /* expval(Observable) / var(Observable): MeasurementsLQubit.hpp:248-295 */
    PrecisionT expval(const Observable<StateVectorT> &ob) {
        const ObsProgram p = ob.program();
        double out = 0;
        check(pl_expval(sv_.handle(), p.prog.data(), p.coeffs.data(), p.mats.data(), &out));
        return static_cast<PrecisionT>(out);
    }
    PrecisionT var(const Observable<StateVectorT> &ob) {
        const ObsProgram p = ob.program();
        double out = 0;
        check(pl_var(sv_.handle(), p.prog.data(), p.coeffs.data(), p.mats.data(), &out));
        return static_cast<PrecisionT>(out);
    }
    /* expval / var by name: MeasurementsLQubit.hpp:167-181, 298-317 */
    PrecisionT expval(const std::string &operation, const std::vector<size_t> &wires) {
        const auto w = StateVectorT::to_int(wires);
        double out = 0;
        check(pl_expval_named(sv_.handle(), gateId(operation), w.data(), static_cast<int>(w.size()), &out));
        return static_cast<PrecisionT>(out);
    }
    PrecisionT var(const std::string &operation, const std::vector<size_t> &wires) {
        const auto w = StateVectorT::to_int(wires);
        double out = 0;
        check(pl_var_named(sv_.handle(), gateId(operation), w.data(), static_cast<int>(w.size()), &out));
        return static_cast<PrecisionT>(out);
    }
    /* expval(matrix, wires): MeasurementsLQubit.hpp:145-158 */
    PrecisionT expval(const std::vector<ComplexT> &matrix, const std::vector<size_t> &wires) {
        std::vector<double> m;
        for (const auto &z : matrix) {
            m.push_back(z.real());
            m.push_back(z.imag());
        }
        const auto w = StateVectorT::to_int(wires);
        double out = 0;
        check(pl_expval_matrix(sv_.handle(), m.data(), w.data(), static_cast<int>(w.size()), &out));
        return static_cast<PrecisionT>(out);
    }
    /* list form: MeasurementsLQubit.hpp:222-246 */
    template <typename op_type>
    std::vector<PrecisionT> expval(const std::vector<op_type> &operations_list,
                                   const std::vector<std::vector<size_t>> &wires_list) {
        if (operations_list.size() != wires_list.size()) {
            throw LightningException("The lengths of the list of operations and wires do not match.");
        }
        std::vector<PrecisionT> out;
        for (size_t i = 0; i < operations_list.size(); i++) {
            out.push_back(expval(operations_list[i], wires_list[i]));
        }
        return out;
    }

    /* generate_samples(num_samples): same output layout as
     * MeasurementsLQubit.hpp:482-558 (row s = sample s, column j = wire j),
     * drawn by inverse-CDF sampling on the device.  The seed comes from
     * std::random_device like the reference unless set_seed() fixed it. */
    std::vector<size_t> generate_samples(size_t num_samples) {
        const size_t n = sv_.getNumQubits();
        std::vector<uint64_t> bits(num_samples * n);
        const uint64_t seed = seeded_ ? seed_ : (uint64_t(std::random_device{}()) << 32) ^ std::random_device{}();
        if (seeded_) {
            seed_ += 0x9E3779B97F4A7C15ull * (num_samples + 1);
        }
        check(pl_generate_samples(sv_.handle(), seed, static_cast<int64_t>(num_samples), bits.data()));
        return {bits.begin(), bits.end()};
    }
    /* basis indices for caller-supplied uniforms (inverse CDF, FP64 prefix) */
    std::vector<int64_t> sample_indices(const std::vector<double> &uniforms) {
        std::vector<int64_t> idx(uniforms.size());
        check(pl_sample(sv_.handle(), uniforms.data(), static_cast<int64_t>(uniforms.size()), idx.data()));
        return idx;
    }
    void set_seed(uint64_t seed) {
        seed_ = seed;
        seeded_ = true;
    }

  private:
    const StateVectorT &sv_;
    uint64_t seed_ = 0;
    bool seeded_ = false;
};

} // namespace Pennylane::LightningAMD
