/*
 * libinvariantq.so — MI355X (gfx950) statevector / density-matrix engine.
 *
 * Drop-in C ABI for the reference InvariantQ native core.  Every name in the
 * first three sections replaces the same name of reference
 * core/include/c_api.h (line numbers cited), with identical C signatures, as
 * bound by invariantq/core/binding.py:54-299.  Semantics follow
 * core/src/api.cpp:14-316 with two documented changes:
 *   - apply_fused uses the MSB-first local index for every k (qubits[0] is the
 *     most significant bit of the row index), and k may be 1..5;
 *   - invalid input never aborts: the call leaves the state unchanged and sets
 *     a thread-local code read with iq_last_error().
 * Matrices are split real/imag, row-major, copied before the call returns.
 * Gate calls are asynchronous on the handle's HIP stream; calls returning data
 * wait for it.  Handles are not thread-safe (as in the reference).
 */
#ifndef INVARIANTQ_H
#define INVARIANTQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------- StateVector (c_api.h:10-24) */
void*  iq_statevector_create(size_t num_qubits);                                                     /* c_api.h:11 */
void   iq_statevector_free(void* sv);                                                                /* c_api.h:12 */
void   iq_statevector_reset(void* sv);                                                               /* c_api.h:13 */
void   iq_statevector_apply_1q(void* sv, size_t target, const double* u_real, const double* u_imag); /* c_api.h:14 */
void   iq_statevector_apply_2q(void* sv, size_t q0, size_t q1, const double* u_real,
                               const double* u_imag);                                                /* c_api.h:15 */
void   iq_statevector_apply_fused(void* sv, const size_t* qubits, size_t k, const double* u_real,
                                  const double* u_imag);                                             /* c_api.h:16 */
int    iq_statevector_measure(void* sv, size_t target, double rand_val);                             /* c_api.h:17 */
void   iq_statevector_reset_qubit(void* sv, size_t target, double rand_val);                         /* c_api.h:18 */
double iq_statevector_measure_prob(void* sv, size_t target, int outcome);                            /* c_api.h:19 */
void   iq_statevector_get_amplitudes(void* sv, double* real_out, double* imag_out);                  /* c_api.h:20 */
void   iq_statevector_get_probabilities(void* sv, double* probs_out);                                /* c_api.h:21 */
void   iq_statevector_sample(void* sv, size_t num_shots, uint64_t seed, uint64_t* results_out);      /* c_api.h:22 */
double iq_statevector_entropy(void* sv, const size_t* qubits, size_t num_subsys_qubits);             /* c_api.h:23 */
/*     entropy: up to 15 subsystem qubits; NaN (error set) when it cannot be computed. */
void   iq_statevector_expval_pauli(void* sv, const char* pauli_str, const size_t* qubits, size_t len,
                                   double* val_real, double* val_imag);                              /* c_api.h:24 */

/* ------------------------------------------------- DensityMatrix (c_api.h:26-38) */
void*  iq_density_matrix_create(size_t num_qubits);                                                  /* c_api.h:27 */
void   iq_density_matrix_free(void* dm);                                                             /* c_api.h:28 */
void   iq_density_matrix_reset(void* dm);                                                            /* c_api.h:29 */
void   iq_density_matrix_apply_1q_unitary(void* dm, size_t target, const double* u_real,
                                          const double* u_imag);                                     /* c_api.h:30 */
void   iq_density_matrix_apply_2q_unitary(void* dm, size_t q0, size_t q1, const double* u_real,
                                          const double* u_imag);                                     /* c_api.h:31 */
void   iq_density_matrix_apply_kraus_1q(void* dm, size_t target, size_t num_kraus, const double* k_real,
                                        const double* k_imag);                                       /* c_api.h:32 */
void   iq_density_matrix_get_elements(void* dm, double* real_out, double* imag_out);                 /* c_api.h:33 */
double iq_density_matrix_trace(void* dm);                                                            /* c_api.h:34 */
double iq_density_matrix_purity(void* dm);                                                           /* c_api.h:35 */
double iq_density_matrix_entropy(void* dm);                                                          /* c_api.h:36 */
int    iq_density_matrix_measure(void* dm, size_t target, double rand_val);                          /* c_api.h:37 */
void   iq_density_matrix_reset_qubit(void* dm, size_t target);                                       /* c_api.h:38 */

/* ------------------- Stabilizer / MPS (c_api.h:40-65): exported, not provided.
 * *_create returns NULL with iq_last_error() == IQ_ERR_UNSUPPORTED; the rest are no-ops. */
void*  iq_stabilizer_create(size_t num_qubits);                                                      /* c_api.h:41 */
void   iq_stabilizer_free(void* stab);                                                               /* c_api.h:42 */
void   iq_stabilizer_reset(void* stab);                                                              /* c_api.h:43 */
void   iq_stabilizer_reset_qubit(void* stab, size_t a, double rand_val);                             /* c_api.h:44 */
void   iq_stabilizer_apply_h(void* stab, size_t a);                                                  /* c_api.h:45 */
void   iq_stabilizer_apply_s(void* stab, size_t a);                                                  /* c_api.h:46 */
void   iq_stabilizer_apply_sdg(void* stab, size_t a);                                                /* c_api.h:47 */
void   iq_stabilizer_apply_cnot(void* stab, size_t a, size_t b);                                     /* c_api.h:48 */
void   iq_stabilizer_apply_cz(void* stab, size_t a, size_t b);                                       /* c_api.h:49 */
void   iq_stabilizer_apply_swap(void* stab, size_t a, size_t b);                                     /* c_api.h:50 */
void   iq_stabilizer_apply_x(void* stab, size_t a);                                                  /* c_api.h:51 */
void   iq_stabilizer_apply_y(void* stab, size_t a);                                                  /* c_api.h:52 */
void   iq_stabilizer_apply_z(void* stab, size_t a);                                                  /* c_api.h:53 */
int    iq_stabilizer_measure(void* stab, size_t a, double rand_val);                                 /* c_api.h:54 */
void   iq_stabilizer_to_string(void* stab, char* buffer, size_t max_len);                            /* c_api.h:55 */
void*  iq_mps_create(size_t num_qubits, size_t max_bond_dim, double cutoff);                         /* c_api.h:58 */
void   iq_mps_free(void* mps);                                                                       /* c_api.h:59 */
void   iq_mps_reset(void* mps);                                                                      /* c_api.h:60 */
void   iq_mps_apply_1q(void* mps, size_t target, const double* u_real, const double* u_imag);        /* c_api.h:61 */
void   iq_mps_apply_2q(void* mps, size_t q0, size_t q1, const double* u_real, const double* u_imag);  /* c_api.h:62 */
double iq_mps_entropy(void* mps, size_t bond_idx);                                                   /* c_api.h:63 */
void   iq_mps_to_statevector(void* mps, double* real_out, double* imag_out);                         /* c_api.h:64 */
int    iq_mps_measure(void* mps, size_t target, double rand_val);                                    /* c_api.h:65 */

/* ------------------------------------------ Additions (no reference counterpart) */
#define IQ_OK 0
#define IQ_ERR_INVALID_ARGUMENT 1
#define IQ_ERR_ZERO_PROBABILITY 2
#define IQ_ERR_OUT_OF_MEMORY 3
#define IQ_ERR_DEVICE 4
#define IQ_ERR_UNSUPPORTED 5

#define IQ_PRECISION_C128 0 /* complex128 amplitudes: the exact mode */
#define IQ_PRECISION_C64 1  /* complex64 amplitudes, FP64 reductions: the fast mode */

/* Error reporting in place of the reference's C++ exceptions
 * (core/src/statevector.cpp:27-29,50-52,100-102,173-175). */
int         iq_last_error(void);
const char* iq_last_error_message(void);
void        iq_clear_error(void);
const char* iq_backend_name(void);
/* Gate kernel selection: 0 automatic, 1 direct (group per lane), 2 LDS-tiled, 3 matrix-core (5-qubit blocks).
 * Automatic also sends diagonal matrices to a diagonal pass and two- or
 * three-qubit blockdiag(I, U) matrices (U on one qubit, every control at bit 4
 * or above) to a pass over the controls-set part of the state; the forced
 * paths apply every matrix densely. */
void        iq_set_gate_path(int path);
/* Matrix-core kernel options. Defaults are per target set; 0 / -1 restores them.
 * - batches_per_iteration: groups-of-16 batches per wave iteration (1, 2, 4). Default:
 *   2 for complex128 sets whose instructions cover whole 128-byte lines and for
 *   complex64 sets with a target at qubit 10 or above, else 4.
 * - matrix_in_lds: matrix fragments in LDS (1) or VGPRs (0; default).
 * - nontemporal: streaming loads and stores (1) or cached (0). Default: on when every
 *   instruction covers whole 128-byte lines. 0 also keeps the direct, diagonal and
 *   whole-state kernels off the hint.
 * - grid_waves: waves in the grid. Default: 524288 for whole-line complex128 sets,
 *   131072 for the other complex128 sets, 262144 for complex64.
 * - batch_rotation: rotation of the super-batch visit order, in bits. Default (from
 *   2^16 super-batches): 1 for whole-line complex128 sets, 2 for the others with a
 *   target at qubit 10 or above, else 0. */
void        iq_set_mfma_options(int batches_per_iteration, int matrix_in_lds, int nontemporal, int grid_waves,
                                int batch_rotation);

void*  iq_statevector_create_precision(size_t num_qubits, int precision);
int    iq_statevector_get_precision(void* sv);
void   iq_statevector_set_amplitudes(void* sv, const double* real_in, const double* imag_in);
/* The amplitudes as interleaved (re, im) FP64 pairs (2 * 2^n doubles, e.g. a NumPy complex128
 * array), for callers that hold complex arrays.  No reference counterpart: get_amplitudes
 * (c_api.h:20) exports split arrays; these copy straight into or out of the caller's array. */
void   iq_statevector_get_amplitudes_interleaved(void* sv, double* out);
void   iq_statevector_set_amplitudes_interleaved(void* sv, const double* in);
double iq_statevector_norm_squared(void* sv);     /* StateVector::norm_squared, statevector.cpp:206-213 */
void   iq_statevector_normalize(void* sv);        /* StateVector::normalize, statevector.cpp:215-224 */
void   iq_statevector_apply_diagonal(void* sv, const size_t* qubits, size_t k, const double* d_real,
                                     const double* d_imag);
/* Sampling with caller-supplied host uniforms in [0,1) (e.g. drawn from a seed
 * tree): each index is the first whose iq_statevector_get_cdf value reaches
 * the uniform, clamped to dim - 1, as for iq_statevector_sample. */
void   iq_statevector_sample_uniforms(void* sv, size_t num_shots, const double* uniforms, uint64_t* results_out);
/* Sampling that allocates nothing.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same indices as iq_statevector_sample_uniforms for the same uniforms, but
 * `uniforms` and `results_out` are device-accessible memory and the call only enqueues work on
 * the handle's stream (iq_statevector_stream): results are valid once that stream reaches them.
 * Its scan scratch is one static 8 MiB device array per GPU, kept until the process ends, so
 * states of at most 2^31 amplitudes are taken (larger: IQ_ERR_UNSUPPORTED); calls that overlap
 * on one GPU (other handles, other threads) share that array and race. */
void   iq_statevector_sample_uniforms_noalloc(void* sv, size_t num_shots, const double* uniforms,
                                              uint64_t* results_out);
/* Inclusive FP64 cumulative distribution used by sampling (dim doubles,
 * non-decreasing; within rounding of the serial running sum). */
void   iq_statevector_get_cdf(void* sv, double* cdf_out);
/* sum_t coeffs[t] <P_t>; term t uses lens[t] consecutive characters of ops and
 * entries of qubits.  per_term (optional) receives Re/Im <P_t> pairs. */
void   iq_statevector_expval_pauli_sum(void* sv, size_t nterms, const double* coeffs, const char* ops,
                                       const size_t* qubits, const size_t* lens, double* out_real,
                                       double* out_imag, double* per_term);
/* A fused gate sequence in one call (no reference counterpart; the output of a fusion pass,
 * invariantq_amd/fusion.py): gate i acts on ks[i] qubits (1..5) taken in order from
 * `qubits`, with 4^ks[i] matrix entries in order from u_real / u_imag, MSB-first as
 * iq_statevector_apply_fused.  All gates are checked before any runs. */
void   iq_statevector_apply_fused_sequence(void* sv, size_t count, const size_t* ks, const size_t* qubits,
                                           const double* u_real, const double* u_imag);
/* Sampling tasks of a distributed job in one call (no reference counterpart; the reference
 * re-simulates per task, invariantq/distributed/worker.py:52-95, with the seeds of
 * invariantq/distributed/planner.py:110-115): task t draws shots[t] indices exactly as
 * iq_statevector_sample(sv, shots[t], seeds[t], ...); results concatenated in task order. */
void   iq_statevector_sample_tasks(void* sv, size_t ntasks, const uint64_t* seeds, const uint64_t* shots,
                                   uint64_t* results_out);
/* Reduced density matrix of `qubits` (k = 1..15; bit b of a row index = qubits[b]), 2^k x 2^k
 * row-major (the reference's StateVector::partial_trace, core/src/statevector.cpp:293-330, a
 * class method with no C export). */
void   iq_statevector_reduced_density_matrix(void* sv, const size_t* qubits, size_t k, double* re_out,
                                             double* im_out);
/* Where the entropy's capped Jacobi runs: on the device from `device_jacobi_min_kept` kept
 * qubits (1..16, 16 = always on the host; 0 restores the default, 10).  No reference
 * counterpart (its Jacobi is serial, core/src/statevector.cpp:333-393); both placements give
 * the same entropy bit for bit. */
void   iq_set_entropy_options(int device_jacobi_min_kept);
void   iq_statevector_synchronize(void* sv);
void*  iq_statevector_device_ptr(void* sv);  /* device address of the amplitudes (re, im interleaved) */
void*  iq_statevector_stream(void* sv);      /* hipStream_t the handle's work is ordered on */

void   iq_density_matrix_set_elements(void* dm, const double* real_in, const double* imag_in);
/* rho as interleaved (re, im) FP64 pairs, row-major (2 * 4^n doubles); no reference counterpart
 * (get_elements, c_api.h:37, exports split arrays). */
void   iq_density_matrix_get_elements_interleaved(void* dm, double* out);
void   iq_density_matrix_set_elements_interleaved(void* dm, const double* in);
/* 1- or 2-qubit Kraus channel (class-only in the reference, density_matrix.cpp:165-191). */
void   iq_density_matrix_apply_kraus(void* dm, const size_t* qubits, size_t k, size_t num_kraus,
                                     const double* k_real, const double* k_imag);
/* <Z_q> for every qubit q from the diagonal (nq doubles). */
/* Caller-composed superoperator on 1 or 2 qubits (gate + channel fusion; no reference
 * counterpart: the reference applies each gate and channel separately,
 * invariantq/simulator/density_matrix.py:35-58).  S is 4^k x 4^k row-major on the local
 * index (row bits of qubits, then column bits), MSB-first, S = sum K (x) conj(K). */
void   iq_density_matrix_apply_superop(void* dm, const size_t* qubits, size_t k, const double* s_real,
                                       const double* s_imag);
void   iq_density_matrix_expval_z(void* dm, double* out);
void   iq_density_matrix_synchronize(void* dm);
void*  iq_density_matrix_stream(void* dm);

/* ------------------------------------------------------- parametric circuits
 * No reference counterpart (the reference's VQE is gradient-free,
 * invariantq/benchmarks/vqe.py:81-98).  A circuit handle owns two device states
 * (psi and the adjoint state) and a gate tape; gates follow
 * invariantq/ir/qir.py:90-136, Pauli strings core/src/statevector.cpp:258-291. */
#define IQ_GATE_RX 0      /* param_index[0] = theta */
#define IQ_GATE_RY 1
#define IQ_GATE_RZ 2
#define IQ_GATE_U3 3      /* param_index[0..2] = theta, phi, lambda */
#define IQ_GATE_CZ 4
#define IQ_GATE_CX 5      /* q0 = control */
#define IQ_GATE_MATRIX1 6 /* fixed 2x2, m_re/m_im row-major */
#define IQ_GATE_MATRIX2 7 /* fixed 4x4, local index 2*bit(q0) + bit(q1) */

void*  iq_circuit_create(size_t num_qubits, int precision);
void   iq_circuit_free(void* c);
int    iq_circuit_add_gate(void* c, int kind, size_t q0, size_t q1, const int64_t* param_index,
                           const double* m_re, const double* m_im);
size_t iq_circuit_num_gates(void* c);
size_t iq_circuit_num_segments(void* c);
/* Observable sum_t coeffs[t] P_t, strings packed as in iq_statevector_expval_pauli_sum. */
int    iq_circuit_set_observable(void* c, size_t nterms, const double* coeffs, const char* ops,
                                 const size_t* qubits, const size_t* lens);
double iq_circuit_expectation(void* c, const double* params, size_t nparams);
/* Adjoint-method <H> and d<H>/dparams (FP64 accumulation on the GPU). */
int    iq_circuit_gradient(void* c, const double* params, size_t nparams, double* energy_out, double* grad_out);
/* Parameter-shift gradient on the same engine for params listed in `which` (all if NULL). */
int    iq_circuit_parameter_shift(void* c, const double* params, size_t nparams, const int64_t* which,
                                  size_t nwhich, double* grad_out);
/* The state after the last evaluation: the circuit's output after an expectation or parameter
 * shift; after a gradient, the output when pass outputs were stored (see below), else |0..0>
 * up to rounding (the reverse sweep takes psi back in place). */
void   iq_circuit_get_amplitudes(void* c, double* real_out, double* imag_out);
/* The forward sweep alone (no observable needed): the circuit's state becomes U(params)|0..0>,
 * asynchronously on iq_circuit_stream(c).  A gate list with no parameters runs with params NULL
 * and nparams 0.  0 on success. */
int    iq_circuit_run(void* c, const double* params, size_t nparams);
/* Copies the circuit's state into a statevector handle of the same qubit count and precision
 * (device to device, after the circuit's work, ordered on the statevector's stream), so the
 * statevector entry points (probabilities, sampling, measurement) act on it.  0 on success. */
int    iq_circuit_copy_state(void* c, void* sv);
void*  iq_circuit_stream(void* c);
/* Device memory a gradient may hold for the forward sweep's pass outputs
 * (default 16 GiB).  While they fit, the reverse sweep reads psi back from them
 * instead of taking it back gate by gate; 0 turns them off.  -1 for a null handle. */
int    iq_circuit_set_stored_budget(void* c, size_t bytes);
/* Reverse passes over stored pass outputs stage lambda alone in LDS and read psi from its
 * slot (1, the default), or stage both (0).  The gradients are the same bit for bit. */
int    iq_circuit_set_lambda_tiles(void* c, int on);
/* Per-launch timing (off by default; a profiling aid): while on, each launch of the tiled
 * sweeps (16 qubits and up) is bracketed by HIP events on the circuit's stream.  After an
 * expectation or gradient, iq_circuit_launch_times waits for the stream and returns the
 * number of timed launches, writing up to `cap` kinds (0 forward pass, 1 observable, 2 reverse
 * pass, 3 two-qubit gate pass, 4 gradient finalize) and durations in ms. */
int    iq_circuit_set_timing(void* c, int on);
size_t iq_circuit_launch_times(void* c, int* kinds, double* ms, size_t cap);

/* ------------------------------------------- batched sweep, one persistent launch
 * No reference counterpart.  Circuits are the ansatz above (RY, RZ on every
 * qubit, CZ ring, per layer) with H = sum Z_q Z_{q+1} + hx sum X_q (open chain). */
void*  iq_batch_create(int precision);
void   iq_batch_free(void* b);
int    iq_batch_max_tile_bits(int precision);
/* Qubits fused per register pass of the tile sweeps (2..4 for complex64, 2..3 for complex128;
 * 0 restores the default, the largest), chosen per batch by the autotuner. */
int    iq_batch_max_group_width(int precision);
int    iq_batch_set_group_width(void* b, int width);
/* params: 2*nq[c]*layers[c] values per circuit, concatenated; grads_out has the same
 * layout.  tile_bits[c] (0 = widest) and block_threads (0 = 512) come from the
 * autotuner; num_blocks 0 = one workgroup per compute unit. */
int    iq_batch_hea_tfim(void* b, size_t ncirc, const int32_t* nq, const int32_t* layers, const double* params,
                         double hx, const int32_t* tile_bits, int block_threads, int num_blocks,
                         double* energies_out, double* grads_out, double* elapsed_ms);

#ifdef __cplusplus
}
#endif

#endif /* INVARIANTQ_H */
