/*
 * pl_amd.h -- C ABI of the MI355X (gfx950) state-vector engine.
 *
 * This is the one boundary between host code and the HIP kernels.  The host
 * C++ classes in lightning_unification_amd/csrc/host/ (StateVector,
 * Measurements, Observables, AdjointJacobian -- the C++ API surface of
 * Lightning-Qubit) and the pybind11 module `lightning_qubit_ops` sit on top of
 * it; a ctypes / cgo / N-API binding can sit on it directly.
 *
 * Conventions
 *  - Every function returns 0 on success.  On failure it returns a nonzero
 *    status and pl_last_error() (thread-local) holds the message:
 *      PL_ERR_INVALID  invalid argument    (reference: std::invalid_argument /
 *                                           PL_ABORT -> LightningException,
 *                                           core/src/utils/Error.hpp:22-128)
 *      PL_ERR_RANGE    unknown name / id   (std::out_of_range from
 *                                           unordered_map::at,
 *                                           DynamicDispatcher.hpp:223-226)
 *      PL_ERR_DEVICE   HIP runtime failure
 *  - A state vector is an opaque handle to 2^n complex amplitudes resident in
 *    HBM, interleaved (re, im), complex64 (PL_C64) or complex128 (PL_C128).
 *    Wire w is bit (n-1-w) of the basis index, as in Lightning-Qubit
 *    (GateImplementationsLM.hpp:209).
 *  - Gate ids follow Pennylane::Gates::GateOperation
 *    (core/src/gates/GateOperation.hpp:27-70): Identity=0 ... MultiRZ=33.
 *    Generator calls take the id of the gate whose generator is applied.
 *  - Parameters always cross as double and matrices as interleaved double
 *    complex, row-major, wires[0] = most significant bit of the matrix index;
 *    they are rounded to the state's precision on the device side.
 *  - Tape encoding (OpsData, core/src/algorithms/JacobianData.hpp:40-200):
 *      nops; ids[nops] (-1 = state preparation, skipped by the adjoint);
 *      nwires[nops] + wires[]; inverse[nops]; nparams[nops] + params[].
 *  - Observable encoding (Observable tree, core/src/observables/Observables.hpp):
 *      int32 program [nobs, { is_ham, nterms, { nfactors,
 *                      { kind, gate_id, nw, w_0 .. w_{nw-1} } } } ]
 *      kind 0 = named gate, 1 = Hermitian matrix (4^nw complex taken in order
 *      from `mats`); is_ham = 1 takes one coefficient per term, in order,
 *      from `coeffs`.
 *  - Calls on one handle are ordered on that handle's HIP stream; calls that
 *    return host data synchronise.  Handles are not thread-safe.
 */
#ifndef PL_AMD_H
#define PL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PL_C64 = 0, PL_C128 = 1 };
enum { PL_OK = 0, PL_ERR_INVALID = 1, PL_ERR_RANGE = 2, PL_ERR_DEVICE = 3 };

typedef struct pl_sv *pl_sv_t;

/* Thread-local message of the last failure on this thread. */
const char *pl_last_error(void);

/* Device name, CU count, HBM size (replaces runtime_info / backend_info,
 * core/src/bindings/Bindings.hpp:203-268, LQubitBindings.hpp:251-264). */
int pl_device_info(char *buf, size_t len);

/* ---- state lifecycle --------------------------------------------------- */

/* New device state |0...0> (StateVectorLQubitManaged(num_qubits),
 * simulators/lightning_qubit/StateVectorLQubitManaged.hpp:72-79). */
int pl_sv_create(int num_qubits, int dtype, pl_sv_t *out);
int pl_sv_destroy(pl_sv_t sv);
int pl_sv_num_qubits(pl_sv_t sv, int *out);
int pl_sv_dtype(pl_sv_t sv, int *out);
/* Host -> device copy of 2^n amplitudes (StateVectorLQubitRaw(ptr, len) /
 * updateData(ptr, len), StateVectorLQubitRaw.hpp:78-87,
 * StateVectorLQubitManaged.hpp:166-169).  nbytes must equal the state size.
 * host may be pageable memory.  Every host array of 8 MiB or more that a call
 * reads or fills (states, probabilities, samples, uniforms, the VJP's
 * cotangent, sparse matrices) is copied through two pinned host slots of
 * 64 MiB that the library allocates at the first such copy and holds until
 * the process ends; concurrent copies through them take turns. */
int pl_sv_upload(pl_sv_t sv, const void *host, size_t nbytes);
/* Device -> host copy (getDataVector / reading the numpy buffer), after the
 * work queued on sv; returns with every byte in host. */
int pl_sv_download(pl_sv_t sv, void *host, size_t nbytes);
/* dst <- src (copy constructor, StateVectorLQubitManaged.hpp:89-94).  The
 * copy is queued on dst's stream after src's queued work, and src's stream
 * waits for it, so work queued on either handle afterwards sees the copy
 * complete. */
int pl_sv_copy(pl_sv_t dst, pl_sv_t src);
/* Bytes moved host -> device by pl_sv_upload and device -> host by
 * pl_sv_download since the library was loaded (either pointer may be NULL).
 * No reference counterpart: Lightning-Qubit's state is host memory.  It lets
 * the drop-in module show that re-wrapping an unchanged, device-mirrored
 * buffer moves no bytes. */
int pl_transfer_counters(uint64_t *h2d_bytes, uint64_t *d2h_bytes);
/* |0...0> in place. */
int pl_sv_reset(pl_sv_t sv);
/* Basis state |index> in place (BasisState preparation). */
int pl_sv_set_basis_state(pl_sv_t sv, uint64_t index);
/* Raw device pointer of the amplitudes (for interop; stays owned by sv). */
int pl_sv_device_ptr(pl_sv_t sv, void **out);
int pl_sv_synchronize(pl_sv_t sv);

/* ---- gates --------------------------------------------------------------- */

/* StateVectorLQubit::applyOperation(name, wires, inverse, params)
 * (StateVectorLQubit.hpp:231-240; kernels GateImplementationsLM.hpp:389-1352). */
int pl_sv_apply_gate(pl_sv_t sv, int gate_id, const int *wires, int nwires,
                     int inverse, const double *params, int nparams);
/* StateVectorBase::applyOperations (StateVectorBase.hpp:84-120): a whole tape
 * in one call. */
int pl_sv_apply_ops(pl_sv_t sv, int nops, const int *ids, const int *nwires,
                    const int *wires, const uint8_t *inverse,
                    const int *nparams, const double *params);
/* Deferred gates (no counterpart in the reference: its applyOperation
 * returns with the gate applied; here the result is the same by the time
 * anything can observe it).  With on != 0, pl_sv_apply_gate and
 * pl_sv_apply_ops check each gate as they otherwise would (same errors, at
 * the call) and queue it; the queue is applied as tile passes, with the
 * precision's default tiles, before any other call on sv reads or changes
 * the state (measurements, adjoint, copies, downloads, pl_sv_device_ptr,
 * pl_sv_synchronize, timing events), and whenever it holds 4096 gates.  A
 * call that overwrites the whole state (pl_sv_upload, pl_sv_set_basis_state,
 * pl_sv_reset, pl_sv_copy into sv) drops it; pl_expval_noalloc, which must
 * not allocate, fails while gates are queued.  on = 0 applies the queue and
 * returns to gate-by-gate application.  Off by default. */
int pl_sv_set_deferred(pl_sv_t sv, int on);
/* Gates queued on sv and not yet applied. */
int pl_sv_deferred_count(pl_sv_t sv, int *out);
/* StateVectorLQubit::applyMatrix(matrix, wires, inverse)
 * (StateVectorLQubit.hpp:329-347; LM.hpp:204-325, PI.hpp:250-290). */
int pl_sv_apply_matrix(pl_sv_t sv, const double *matrix, const int *wires,
                       int nwires, int inverse);
/* StateVectorLQubit::applyGenerator(name, wires, adj) -> scale
 * (StateVectorLQubit.hpp:266-275; LM.hpp:1355-1640). */
int pl_sv_apply_generator(pl_sv_t sv, int gate_id, const int *wires,
                          int nwires, int adj, double *scale);

/* Observable::applyInPlace (core/src/observables/Observables.hpp:98-324,
 * ObservablesLQubit.hpp:351-355 for Hamiltonians): sv <- O sv for the one
 * observable held by prog/coeffs/mats. */
int pl_sv_apply_observable(pl_sv_t sv, const int *prog, const double *coeffs,
                           const double *mats);

/* ---- measurements ---------------------------------------------------------- */

/* Measurements::expval(Observable) / var(Observable)
 * (MeasurementsLQubit.hpp:248-295).  prog/coeffs/mats hold one observable. */
int pl_expval(pl_sv_t sv, const int *prog, const double *coeffs,
              const double *mats, double *out);
int pl_var(pl_sv_t sv, const int *prog, const double *coeffs,
           const double *mats, double *out);
/* pl_expval without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Takes Pauli words and Hamiltonians of Pauli words (named PauliX / PauliY /
 * PauliZ / Identity factors) and fails with PL_ERR_INVALID otherwise.  Its
 * term tables and partial sums live in a fixed 2 MiB device array per GPU
 * (held until the process ends); an observable needing more fails with
 * PL_ERR_INVALID (a 100-term Hamiltonian needs about 0.2 MiB), and so does a
 * call while deferred gates are queued on sv (applying them may allocate;
 * pl_sv_synchronize applies them first).  Calls that overlap on one GPU
 * share that array and race. */
int pl_expval_noalloc(pl_sv_t sv, const int *prog, const double *coeffs,
                      const double *mats, double *out);
/* Measurements::expval/var(name, wires) (MeasurementsLQubit.hpp:167-181,298-317). */
int pl_expval_named(pl_sv_t sv, int gate_id, const int *wires, int nwires,
                    double *out);
int pl_var_named(pl_sv_t sv, int gate_id, const int *wires, int nwires,
                 double *out);
/* Measurements::expval(matrix, wires) (MeasurementsLQubit.hpp:145-158). */
int pl_expval_matrix(pl_sv_t sv, const double *matrix, const int *wires,
                     int nwires, double *out);
/* Measurements::probs() (nwires < 0, out has 2^n entries) and probs(wires)
 * (out has 2^nwires entries), MeasurementsLQubit.hpp:79-135.  As in the
 * reference, output axis j (most significant first) is wire
 * sorted(wires)[argsort(wires)[j]]: the caller's order whenever that order's
 * permutation is its own inverse (sorted, reversed, one swap). */
int pl_probs(pl_sv_t sv, const int *wires, int nwires, double *out);
/* Inverse-CDF sampling with caller-supplied uniforms in [0,1): idx[s] is the
 * basis index with C_i < u_s <= C_{i+1} (C exclusive prefix of |psi|^2 in
 * FP64, C_N := 1), the rule of lightning_kokkos
 * MeasurementsKokkos.hpp:700-741 / MeasuresFunctors.hpp:55-105.  Replaces the
 * unseedable alias sampler of MeasurementsLQubit.hpp:482-558. */
int pl_sample(pl_sv_t sv, const double *uniforms, int64_t nsamples,
              int64_t *idx);
/* Same draws expanded to bits, samples[s*n + (n-1-j)] = (idx_s >> j) & 1
 * (the layout of generate_samples, MeasurementsLQubit.hpp:546-552); the
 * uniforms come from a SplitMix64 stream seeded by `seed`. */
int pl_generate_samples(pl_sv_t sv, uint64_t seed, int64_t nsamples,
                        uint64_t *bits);

/* Measurements::expval / var of a sparse Hamiltonian in CSR form
 * (MeasurementsLQubit.hpp:183-214, 437-471; SparseLinAlg.hpp): row_map has
 * 2^n + 1 entries (row_map[i] = non-zeros before row i), entries and values
 * (interleaved complex) have nnz each.  One fused pass computes <psi|H psi>
 * and ||H psi||^2 with FP64 accumulation; var = ||H psi||^2 - <psi|H psi>^2.
 * row_map[0] = 0 and row_map[2^n] = nnz are checked on the host; the order of
 * row_map and the range of every column index are checked by the pass, which
 * reads through neither, and the call then fails with PL_ERR_INVALID. */
int pl_expval_sparse(pl_sv_t sv, const int64_t *row_map, int64_t row_map_size,
                     const int64_t *entries, const double *values, int64_t nnz,
                     double *out);
int pl_var_sparse(pl_sv_t sv, const int64_t *row_map, int64_t row_map_size,
                  const int64_t *entries, const double *values, int64_t nnz,
                  double *out);
/* Measurements::generate_samples_metropolis (MeasurementsLQubit.hpp:388-434,
 * 613-640; TransitionKernels.hpp): Metropolis-Hastings samples of |psi|^2 in
 * the generate_samples bit layout.  kernel 0 proposes uniformly among the
 * indices with |psi_i| > epsilon of the state's precision; kernel 1 flips one
 * uniformly chosen bit to a uniformly chosen value.  (The reference's
 * kernel_factory builds kernel 0 for the name "Local" and kernel 1 for
 * "NonZeroRandom"; the pybind layer keeps that mapping.)  nchains independent
 * chains (0 = automatic: one per shot up to 16384) each start from an
 * inverse-CDF draw of |psi|^2 (the reference's one chain starts at index 0;
 * a chain started in the stationary distribution needs no burn-in to be
 * unbiased, which short chains rely on), discard num_burnin steps and record
 * consecutive states, filling the output in chain order; SplitMix64 streams
 * keyed by seed.  The reference's single chain is unseedable, so results
 * agree in distribution only. */
int pl_generate_mcmc_samples(pl_sv_t sv, int kernel, int64_t num_burnin,
                             int64_t nshots, uint64_t seed, int nchains,
                             uint64_t *bits);

/* ---- adjoint Jacobian ------------------------------------------------------ */

/* AdjointJacobian::adjointJacobian(jac, JacobianData{ops, obs, tp})
 * (AdjointJacobianLQubit.hpp:212-336).  The sweep starts from the state held
 * by sv (not modified).  jac is observable-major: jac[k*ntp + t]. */
int pl_adjoint_jacobian(pl_sv_t sv, int nops, const int *ids,
                        const int *nwires, const int *wires,
                        const uint8_t *inverse, const int *nparams,
                        const double *params, const int *prog,
                        const double *coeffs, const double *mats, int ntp,
                        const int *tp, double *jac);

/* ---- precision path: gate fusion ------------------------------------------- */

enum { PL_FUSE_RENORM = 1, PL_FUSE_NO_MFMA = 2 };

/* The operations of StateVectorBase::applyOperations (StateVectorBase.hpp:84-120),
 * multiplied into dense blocks of at most max_wires (1..5) wires first.  A
 * block grows from the first operation not yet placed by scanning ahead in
 * tape order: an operation joins when its wires fit and it shares no wire
 * with an operation the scan passed over, so operations only move ahead of
 * operations on disjoint wires, with which they commute.
 * On complex64 states with n >= 12 every block is widened to 5 wires and
 * applied on FP32 matrix cores; elsewhere blocks use the dense-matrix kernels.
 * flags: PL_FUSE_RENORM renormalises to unit norm (FP64 norm; on the matrix-
 * core path after every block), PL_FUSE_NO_MFMA forces the dense-matrix
 * kernels.  Operations wider than max_wires are applied unfused.
 * *nblocks (may be NULL) receives the number of blocks. */
int pl_sv_apply_ops_fused(pl_sv_t sv, int nops, const int *ids, const int *nwires,
                          const int *wires, const uint8_t *inverse,
                          const int *nparams, const double *params, int max_wires,
                          int flags, int *nblocks);
/* The fusion plan of pl_sv_apply_ops_fused, computed on the host only (no
 * device is touched): order[] lists the tape indices of all operations
 * (state preparation excluded) in application order, block after block; for
 * block b, first_op[b] is the offset of its first operation in order[],
 * block_nops[b] its operation count, nbits[b] its width (0 = one operation
 * applied unfused), bits[8 b ..] ascending bit positions, and
 * mats[2 * 1024 b ..] the 2^nbits x 2^nbits row-major matrix (interleaved
 * double complex) in the block-local index whose bit m is bit position
 * bits[8 b + m].  Blocks narrower than pad_to are widened with identity bits.
 * Fails with PL_ERR_INVALID, and sets *nblocks, when capacity is too small. */
int pl_fuse_tape(int num_qubits, int nops, const int *ids, const int *nwires,
                 const int *wires, const uint8_t *inverse, const int *nparams,
                 const double *params, int max_wires, int pad_to, int capacity,
                 int *nblocks, int *first_op, int *block_nops, int *order,
                 int *nbits, int *bits, double *mats);
/* StateVectorBase::applyOperations (StateVectorBase.hpp:84-120) executed as
 * tile passes: the tape is split into passes, each of which loads tiles of
 * 2^tile_bits amplitudes (the low_bits lowest bit positions plus
 * tile_bits - low_bits higher ones) into the registers of one wavefront,
 * applies every operation of the pass there and writes the tiles back, so the
 * state is read and written once per pass.  Operations move ahead of earlier
 * ones only across operations on disjoint wires, with which they commute.
 * Operations a pass cannot hold are applied by the per-gate kernels.
 * tile_bits in [9, 12] ([9, 11] for complex128: a 2^12-amplitude tile does
 * not fit a wavefront's registers) and <= the qubit count; low_bits in
 * [0, 6].  tile_bits = 0 and low_bits = -1 pick the
 * defaults of the precision (2^11 amplitudes and 3 low bits for complex128,
 * 2^12 and 4 for complex64, clamped to the state); with tile_bits = 0 a state
 * of fewer than 9 qubits runs gate by gate.  flags: PL_FUSE_RENORM rescales
 * the state to unit norm at the end, the norm accumulated in FP64 (the
 * precision form's renormalisation for complex64 states).  *npasses (may be
 * NULL) receives the number of launches. */
int pl_sv_apply_ops_tiled(pl_sv_t sv, int nops, const int *ids, const int *nwires,
                          const int *wires, const uint8_t *inverse,
                          const int *nparams, const double *params, int tile_bits,
                          int low_bits, int flags, int *npasses);
/* The pass plan of pl_sv_apply_ops_tiled, computed on the host only: for
 * pass p, pass_nops[p] operations, pass_tiled[p] (0 = one operation applied
 * by the per-gate kernels), pass_bits[16 p ..] the higher tile bit positions
 * (ascending, -1 padded); order[] lists the tape indices of all operations
 * (state preparation excluded) in application order, pass after pass.
 * Fails with PL_ERR_INVALID, and sets *npasses, when capacity is too small. */
int pl_tile_plan(int num_qubits, int nops, const int *ids, const int *nwires,
                 const int *wires, const uint8_t *inverse, const int *nparams,
                 const double *params, int tile_bits, int low_bits, int capacity,
                 int *npasses, int *pass_nops, int *pass_tiled, int *pass_bits,
                 int *order);
/* <psi|psi> accumulated in FP64 (innerProdC(psi, psi), LinearAlgebra.hpp:193-213). */
int pl_sv_norm2(pl_sv_t sv, double *out);
/* ||a - b||_2 accumulated in FP64; a and b must have the same qubit count and
 * may differ in precision (the error of a complex64 run against complex128). */
int pl_sv_distance(pl_sv_t a, pl_sv_t b, double *out);

/* VectorJacobianProduct::operator()(jac, JacobianData{ops, tp}, dy)
 * (VectorJacobianProduct.hpp:77-155): w_t = <dy| d psi / d theta_t> for the
 * trainable parameters tp (sorted), starting from the final state held by sv
 * (not modified).  dy: 2^n interleaved complex; vjp: ntp interleaved complex. */
int pl_vjp(pl_sv_t sv, int nops, const int *ids, const int *nwires,
           const int *wires, const uint8_t *inverse, const int *nparams,
           const double *params, const double *dy, int ntp, const int *tp,
           double *vjp);

/* ---- batched layered circuits ------------------------------------------------ */

/* Many independent circuits of uneven size, forward pass + energy + adjoint
 * gradients in one replayed hipGraph.  No single reference function: it
 * batches what AdjointJacobianLQubit.hpp:212-336 computes per circuit for one
 * observable (K = 1), for the layered-ansatz family
 *   for each of depth[j] layers: for w in 0..n-1: the rotations named by
 *   `pattern` (e.g. "YZ" = RY then RZ, each with its own angle), then the
 *   CNOT entangler (0 none, 1 ladder i->i+1, 2 ring i->(i+1) mod n).
 * prog/coeffs: ncirc observables in the program encoding (Pauli words or
 * Hamiltonians of Pauli words).  tile_width 0 lets the autotuner choose.
 * Angles and gradients of circuit j are laid out layer-major, then wire, then
 * pattern letter -- the tape order of the reference's operations. */
typedef struct pl_batch *pl_batch_t;
int pl_batch_create(int dtype, int ncirc, const int *nqubits, const int *depth,
                    const char *pattern, int entangler, const int *prog,
                    const double *coeffs, int tile_width, pl_batch_t *out);
/* energies[ncirc]; grads[total angles] (may be NULL) */
int pl_batch_run(pl_batch_t b, const double *params, double *energies,
                 double *grads);
int pl_batch_info(pl_batch_t b, int *tile_width, int *launches,
                  int64_t *nparams);
int pl_batch_destroy(pl_batch_t b);

/* ---- learned surrogate -------------------------------------------------------- */

/* A d0 -> d1 -> d2 -> d3 MLP (GELU hidden layers) trained and evaluated on the
 * GPU with BF16 matrix cores and FP32 accumulation, FP32 master weights and
 * Adam.  No reference counterpart: it screens optimiser candidates for the
 * adjoint path (the exact computation of AdjointJacobianLQubit.hpp:212-336
 * verifies every accepted step).  Parameters cross as FP32 in the order W1 [d0 x d1]
 * row-major, b1, W2 [d1 x d2], b2, W3 [d2 x d3], b3; features as FP32 rows of
 * d0 (rounded to BF16 on the device), targets and outputs as FP32 rows of d3.
 * The loss is sum_j w_j (y_j - t_j)^2 averaged over the batch (w_j = 1 unless
 * set). */
typedef struct pl_mlp *pl_mlp_t;
int pl_mlp_create(int d0, int d1, int d2, int d3, int max_batch, uint64_t seed,
                  pl_mlp_t *out);
int pl_mlp_destroy(pl_mlp_t m);
int pl_mlp_num_params(pl_mlp_t m, int64_t *out);
int pl_mlp_get_params(pl_mlp_t m, float *params);
int pl_mlp_set_params(pl_mlp_t m, const float *params); /* also resets Adam */
int pl_mlp_set_loss_weights(pl_mlp_t m, const float *w); /* d3 weights */
/* y[n x d3] for features x[n x d0], n <= max_batch */
int pl_mlp_predict(pl_mlp_t m, const float *x, int n, float *y);
/* loss and its gradient (parameter layout) on one batch, n <= max_batch */
int pl_mlp_gradients(pl_mlp_t m, const float *x, const float *t, int n,
                     double *loss, float *grads);
/* copy a training set to the device (features stored as BF16) */
int pl_mlp_load_data(pl_mlp_t m, const float *x, const float *t, int64_t n);
/* `steps` Adam steps over consecutive minibatches of the loaded data;
 * *mean_loss (may be NULL) is the mean minibatch loss */
int pl_mlp_train(pl_mlp_t m, int64_t steps, int batch, float lr, double *mean_loss);

/* ---- timing ------------------------------------------------------------------ */

/* HIP events recorded on the handle's stream around the calls in between;
 * pl_timer_stop synchronises and returns the elapsed milliseconds. */
int pl_timer_start(pl_sv_t sv);
int pl_timer_stop(pl_sv_t sv, double *ms);
/* Event slots (0 <= slot < 65536) recorded on the handle's stream without
 * synchronising, for per-launch timing; pl_event_elapsed waits for slot b. */
int pl_event_record(pl_sv_t sv, int slot);
int pl_event_elapsed(pl_sv_t sv, int slot_a, int slot_b, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* PL_AMD_H */
