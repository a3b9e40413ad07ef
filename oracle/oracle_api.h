/*
 * Checker-side C interface shared by the two CPU oracles:
 *   - liblq_oracle.so   : this repository's restatement of Lightning-Qubit
 *                         (oracle/lq_oracle.cpp), and
 *   - _ref/libref_lq.so : a thin driver (oracle/ref_driver.cpp) over the
 *                         reference's own LQ sources, compiled in place.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product package links, loads or
 * calls this interface; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do.
 *
 * Both libraries export the same symbols so a test can run one workload on
 * each and compare.  State vectors are interleaved (re, im) arrays of
 * 2^n complex numbers; suffix _c128 takes double, _c64 takes float.
 * Wire 0 is the most significant bit of the basis index
 * (simulators/lightning_qubit/gates/cpu_kernels/GateImplementationsLM.hpp:209).
 *
 * Encodings (identical to the ones the product C-ABI in include/pl_amd.h uses):
 *
 *  gate ids      : order of Pennylane::Gates::GateOperation
 *                  (core/src/gates/GateOperation.hpp:27-70), Identity=0 ...
 *                  MultiRZ=33.  Id -1 marks a state-preparation op
 *                  ("QubitStateVector"/"BasisState") that the adjoint sweep
 *                  skips (AdjointJacobianLQubit.hpp:287-290).
 *  generator ids : the gate id of the parametrised gate whose generator is
 *                  applied (the name-based lookup of
 *                  DynamicDispatcher::applyGenerator, DynamicDispatcher.hpp:478-515).
 *  tape          : nops; ids[nops]; nwires[nops] + wires[]; inverse[nops];
 *                  nparams[nops] + params[] (double, cast to the precision).
 *  observables   : int32 program
 *                    [nobs, { is_ham, nterms, { nfactors,
 *                               { kind, op_id, nw, w_0 .. w_{nw-1} } } } ]
 *                  kind 0 = named gate (NamedObs), 1 = Hermitian matrix
 *                  (HermitianObs, matrix taken in order from `mats`,
 *                  4^nw complex each).  is_ham = 1 builds a Hamiltonian whose
 *                  term coefficients are taken in order from `coeffs`;
 *                  is_ham = 0 must have exactly one term, built as the bare
 *                  factor (nfactors = 1) or a TensorProdObs (nfactors > 1).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *orc_last_error(void);
int orc_num_threads(void);

#define ORC_DECL(SUF, REAL)                                                   \
    int orc_apply_ops_##SUF(REAL *sv, int n, int nops, const int *ids,        \
                            const int *nwires, const int *wires,              \
                            const uint8_t *inverse, const int *nparams,       \
                            const double *params);                            \
    int orc_apply_matrix_##SUF(REAL *sv, int n, const REAL *mat, int nw,      \
                               const int *wires, int inverse);                \
    int orc_apply_generator_##SUF(REAL *sv, int n, int gate_id, int nw,       \
                                  const int *wires, int adj, double *scale);  \
    int orc_expval_##SUF(const REAL *sv, int n, const int *prog,              \
                         const double *coeffs, const REAL *mats,              \
                         double *out);                                        \
    int orc_var_##SUF(const REAL *sv, int n, const int *prog,                 \
                      const double *coeffs, const REAL *mats, double *out);   \
    int orc_expval_named_##SUF(const REAL *sv, int n, int gate_id, int nw,    \
                               const int *wires, double *out);                \
    int orc_var_named_##SUF(const REAL *sv, int n, int gate_id, int nw,       \
                            const int *wires, double *out);                   \
    int orc_expval_matrix_##SUF(const REAL *sv, int n, const REAL *mat,       \
                                int nw, const int *wires, double *out);       \
    int orc_probs_##SUF(const REAL *sv, int n, int nw, const int *wires,      \
                        REAL *out);                                           \
    int orc_adjoint_##SUF(const REAL *sv, int n, int nops, const int *ids,    \
                          const int *nwires, const int *wires,                \
                          const uint8_t *inverse, const int *nparams,         \
                          const double *params, const int *prog,              \
                          const double *coeffs, const REAL *mats, int ntp,    \
                          const int *tp, REAL *jac);                          \
    int orc_vjp_##SUF(const REAL *sv, int n, int nops, const int *ids,        \
                      const int *nwires, const int *wires,                    \
                      const uint8_t *inverse, const int *nparams,             \
                      const double *params, const REAL *dy, int ntp,          \
                      const int *tp, REAL *vjp);                              \
    int orc_expval_sparse_##SUF(const REAL *sv, int n, const int64_t *row_map,\
                                int64_t row_map_size, const int64_t *entries, \
                                const REAL *values, int64_t nnz, double *out);\
    int orc_var_sparse_##SUF(const REAL *sv, int n, const int64_t *row_map,   \
                             int64_t row_map_size, const int64_t *entries,    \
                             const REAL *values, int64_t nnz, double *out);

ORC_DECL(c128, double)
ORC_DECL(c64, float)

/* CDF sampling with injected uniforms (restated in both libraries; the
 * reference's LQ sampler is an unseedable alias method, so only the oracle
 * restatement of the Kokkos CDF rule exists).  idx[s] is the basis index. */
int orc_sample_cdf_c128(const double *sv, int n, int64_t m,
                        const double *uniforms, int64_t *idx);
int orc_sample_cdf_c64(const float *sv, int n, int64_t m,
                       const double *uniforms, int64_t *idx);

#ifdef __cplusplus
}
#endif
