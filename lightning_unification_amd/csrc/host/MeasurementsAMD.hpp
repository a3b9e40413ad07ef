// Measurements over the engine (core/src/measurements/MeasurementsBase.hpp:42-106,
// simulators/lightning_qubit/measurements/MeasurementsLQubit.hpp:60-641).
#pragma once

#include <cstdint>
#include <random>
#include <string>
#include <type_traits>
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
    /* expval / var of a CSR sparse Hamiltonian: MeasurementsLQubit.hpp:183-214, 437-471 */
    template <class index_type>
    PrecisionT expval(const index_type *row_map_ptr, const index_type row_map_size,
                      const index_type *entries_ptr, const ComplexT *values_ptr, const index_type numNNZ) {
        return sparse(pl_expval_sparse, row_map_ptr, row_map_size, entries_ptr, values_ptr, numNNZ);
    }
    template <class index_type>
    PrecisionT var(const index_type *row_map_ptr, const index_type row_map_size,
                   const index_type *entries_ptr, const ComplexT *values_ptr, const index_type numNNZ) {
        return sparse(pl_var_sparse, row_map_ptr, row_map_size, entries_ptr, values_ptr, numNNZ);
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
        check(pl_generate_samples(sv_.handle(), next_seed(num_samples), static_cast<int64_t>(num_samples),
                                  bits.data()));
        return {bits.begin(), bits.end()};
    }
    /* basis indices for caller-supplied uniforms (inverse CDF, FP64 prefix) */
    std::vector<int64_t> sample_indices(const std::vector<double> &uniforms) {
        std::vector<int64_t> idx(uniforms.size());
        check(pl_sample(sv_.handle(), uniforms.data(), static_cast<int64_t>(uniforms.size()), idx.data()));
        return idx;
    }
    /* generate_samples_metropolis(kernelname, num_burnin, num_samples):
     * MeasurementsLQubit.hpp:388-434, with the reference's name mapping:
     * "NonZeroRandom" selects TransitionKernelType::NonZeroRandom, every
     * other name TransitionKernelType::Local (:396-400), and kernel_factory
     * (TransitionKernels.hpp:149-159) builds single-bit flips for the former
     * and proposals uniform over the non-zero amplitudes for the latter. */
    std::vector<size_t> generate_samples_metropolis(const std::string &kernelname, size_t num_burnin,
                                                    size_t num_samples) {
        const size_t n = sv_.getNumQubits();
        std::vector<uint64_t> bits(num_samples * n);
        const int kind = kernelname == "NonZeroRandom" ? 1 : 0;
        check(pl_generate_mcmc_samples(sv_.handle(), kind, static_cast<int64_t>(num_burnin),
                                       static_cast<int64_t>(num_samples), next_seed(num_samples), 0,
                                       bits.data()));
        return {bits.begin(), bits.end()};
    }
    void set_seed(uint64_t seed) {
        seed_ = seed;
        seeded_ = true;
    }

  private:
    uint64_t next_seed(size_t num_samples) {
        const uint64_t seed = seeded_ ? seed_ : (uint64_t(std::random_device{}()) << 32) ^ std::random_device{}();
        if (seeded_) {
            seed_ += 0x9E3779B97F4A7C15ull * (num_samples + 1);
        }
        return seed;
    }
    /* The C-ABI takes int64 indices and interleaved double values.  When the
     * caller's arrays already have that layout (complex128 values,
     * 64-bit indices; std::complex<double> is a pair of doubles) they are
     * passed through as they are; otherwise they are converted once. */
    template <class F, class index_type>
    PrecisionT sparse(F fn, const index_type *row_map_ptr, const index_type row_map_size,
                      const index_type *entries_ptr, const ComplexT *values_ptr, const index_type numNNZ) {
        double out = 0;
        if constexpr (std::is_same_v<PrecisionT, double> && std::is_integral_v<index_type> &&
                      sizeof(index_type) == sizeof(int64_t)) {
            static const int64_t no_entry = 0;
            static const double no_value[2] = {0, 0};
            const auto *en = numNNZ > 0 ? reinterpret_cast<const int64_t *>(entries_ptr) : &no_entry;
            const auto *va = numNNZ > 0 ? reinterpret_cast<const double *>(values_ptr) : no_value;
            check(fn(sv_.handle(), reinterpret_cast<const int64_t *>(row_map_ptr),
                     static_cast<int64_t>(row_map_size), en, va, static_cast<int64_t>(numNNZ), &out));
        } else {
            std::vector<int64_t> rm(row_map_ptr, row_map_ptr + row_map_size),
                en(entries_ptr, entries_ptr + numNNZ);
            std::vector<double> va(2 * static_cast<size_t>(numNNZ) + 2);
            for (index_type j = 0; j < numNNZ; j++) {
                va[2 * j] = values_ptr[j].real();
                va[2 * j + 1] = values_ptr[j].imag();
            }
            en.push_back(0);
            check(fn(sv_.handle(), rm.data(), static_cast<int64_t>(rm.size()), en.data(), va.data(),
                     static_cast<int64_t>(numNNZ), &out));
        }
        return static_cast<PrecisionT>(out);
    }

    const StateVectorT &sv_;
    uint64_t seed_ = 0;
    bool seeded_ = false;
};

} // namespace Pennylane::LightningAMD
