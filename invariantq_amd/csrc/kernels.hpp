// Host-side launch interface of the statevector / density-matrix kernels.
#pragma once

#include "common.hpp"
#include "layout.hpp"

namespace iqg {

// Matrix layout consumed by the gate kernels for one 2^K x 2^K block:
//   [A | B]      (4-multiply form: A = Re U, B = Im U), or
//   [A | B | S]  (3-multiply form, S = A + B),
// each D x D row-major in the state's real type.
size_t gate_matrix_bytes(int prec, int K, bool three_mult);
void pack_gate_matrix(int prec, int K, bool three_mult, const double* re, const double* im, void* out);
bool use_three_mult(int K);
// Matrix-core kernel options: batches of 16 groups per loop iteration (1, 2, 4;
// 0 = default), matrix fragments in LDS (1) or VGPRs (0; -1 = default),
// non-temporal loads/stores (1) or cached (0; -1 = default), waves in the grid (0 = default).
void set_mfma_options(int batches_per_iteration, int matrix_in_lds, int nontemporal, int grid_waves,
                      int batch_rotation);
// Fragment-ordered copy of a 32x32 block for the matrix-core kernel (depends on the targets).
size_t mfma_fragment_bytes(int prec);
void pack_mfma_fragments(int prec, int n, const uint32_t* q, const double* re, const double* im, void* out);

enum GatePath : int { kPathAuto = 0, kPathDirect = 1, kPathTiled = 2, kPathMfma = 3 };

struct GateDesc {
    int K = 0;                 // number of target qubits, 1..5
    uint32_t q[kMaxK] = {};    // target list, MSB-first: q[0] is the most significant local bit
    const double* re = nullptr;  // host row-major D x D (kept for by-value launches, K <= 3)
    const double* im = nullptr;
    const void* dmat = nullptr;  // device copy: pack_gate_matrix layout, or pack_mfma_fragments for the MFMA path
    int path = kPathAuto;
};

// psi <- U psi on the K target qubits.  Returns the launch status.
hipError_t launch_gate(void* psi, int n, int prec, hipStream_t st, const GateDesc& g);
// Diagonal gate: psi[i] *= d[local(i)] (d host, 2^K entries, MSB-first local index).
hipError_t launch_diag(void* psi, int n, int prec, hipStream_t st, int K, const uint32_t* q,
                       const double* dre, const double* dim);
// Controlled one-qubit gate on target t (nc = 1 or 2 controls c): psi <- U psi
// on the amplitudes with every control bit set; ure/uim row-major 2x2 (host).
hipError_t launch_controlled(void* psi, int n, int prec, hipStream_t st, const uint32_t* c, int nc, uint32_t t,
                             const double* ure, const double* uim);
int choose_path(int n, int prec, const GateDesc& g);
// Streaming (non-temporal) hints in whole-state passes: on unless turned off
// through the matrix-core options (nontemporal = 0).
bool stream_hint_enabled();

// State initialisation and element-wise passes.
hipError_t launch_init_basis(void* psi, int n, int prec, uint64_t index, hipStream_t st);
// Copies `bytes` (a multiple of 16) from device-accessible pinned host memory.
hipError_t launch_upload(void* dst, const void* src_host, size_t bytes, hipStream_t st);
hipError_t launch_collapse(void* psi, int n, int prec, int target, int outcome, double factor, hipStream_t st);
hipError_t launch_scale(void* psi, uint64_t count, int prec, double factor, hipStream_t st);
hipError_t launch_split(const void* psi, uint64_t first, uint64_t count, int prec, double* re, double* im, hipStream_t st);
hipError_t launch_join(void* psi, uint64_t first, uint64_t count, int prec, const double* re, const double* im, hipStream_t st);
// Amplitudes [first, first + count) as interleaved FP64 (re, im) pairs, and back.
hipError_t launch_interleave(const void* psi, uint64_t first, uint64_t count, int prec, double* out, hipStream_t st);
hipError_t launch_deinterleave(void* psi, uint64_t first, uint64_t count, int prec, const double* in, hipStream_t st);
hipError_t launch_probs(const void* psi, uint64_t count, int prec, double* out, hipStream_t st);

// Deterministic FP64 reductions.  Each writes `m` values into `out` (device),
// using `part` (device, kRedBlocks * m doubles) as per-block partial sums.
// Sum `nblocks` per-block partials of m values each (layout part[b * m + j]) into out[0..m).
hipError_t launch_finalize(const double* part, unsigned nblocks, int m, double* out, hipStream_t st);
hipError_t reduce_norm(const void* psi, uint64_t count, int prec, double* part, double* out, hipStream_t st);
hipError_t reduce_bit_probs(const void* psi, int n, int prec, int target, double* part, double* out, hipStream_t st);
// Pauli strings sharing one X/Y flip mask: out[2s], out[2s+1] = Re, Im <psi|P_s|psi>.
hipError_t reduce_pauli_group(const void* psi, int n, int prec, uint64_t xmask, int nstr, const uint64_t* zy_masks,
                              const int* ny, double* part, double* out, hipStream_t st);

// Density matrix (vectorised rho, row-major: element (r, c) is index r * 2^nq + c).
// mode 0: trace; mode 1: <Z_q> for every q (nq values); mode 2: P(bit t = 0), P(bit t = 1) clamped later;
// mode 3: diagonal Shannon entropy -sum p log2 p over p > 1e-14.
hipError_t reduce_dm_diag(const void* rho, int nq, int mode, int target, double* part, double* out, hipStream_t st);
hipError_t launch_dm_collapse(void* rho, int nq, int target, int outcome, double factor, hipStream_t st);

// Sampling: the first index whose FP64 CDF value of |psi_i|^2 reaches each
// uniform, clamped to count - 1.  scratch holds scan_scratch_doubles(count)
// doubles (chunk offsets and running maxima).
size_t scan_scratch_doubles(uint64_t count);
// true when per-chunk searches (launch_cdf_chunks + launch_sample_chunks)
// cost less than writing the CDF (launch_cdf + launch_search): up to twice
// as many shots as 4096-amplitude chunks.  Both give the same indices.
bool sample_without_cdf(uint64_t count, uint64_t shots);
hipError_t launch_cdf_chunks(const void* psi, uint64_t count, int prec, double* scratch, hipStream_t st);
hipError_t launch_sample_chunks(const void* psi, uint64_t count, int prec, const double* scratch, const double* u,
                                uint64_t shots, uint64_t* out, hipStream_t st);
// Inclusive FP64 CDF (monotone) into cdf (count doubles).
hipError_t launch_cdf(const void* psi, uint64_t count, int prec, double* cdf, double* scratch, hipStream_t st);
hipError_t launch_search(const double* cdf, uint64_t count, const double* u, uint64_t shots, uint64_t* out, hipStream_t st);

// Reduced density matrix of the kept qubits (bit b of a row index = qubit
// keep[b]), d x d row-major, exactly Hermitian.  `scratch` (scratch_elems
// double2) holds the workgroups' partial sums, and for 4 or more kept qubits
// the state reordered to (environment, kept) (2^n double2) unless keep is
// 0..k-1 in order on a complex128 state.  1-3 kept: the environment sum over
// many workgroups; 4 or more (and 2 or more traced): the FP64 matrix cores,
// split over environment chunks when the matrix alone gives few workgroups;
// otherwise one thread per element.
hipError_t launch_partial_trace(const void* psi, int n, int prec, const uint32_t* keep, int k, double2* out,
                                double2* scratch, uint64_t scratch_elems, hipStream_t st);
// The capped Jacobi's device side (jacobi_rows.hpp) on a d x d Hermitian
// matrix A (row-major double2):
// rotate: the rotation (p, q) with cosine c and sine (sr, si), A(p, p) = npp,
//   A(q, q) = nqq, A(p, q) = A(q, p) = 0; colp[i], colq[i] = the new A(i, p),
//   A(i, q) for i < ncol; zeroes *extra_n for the scan that follows.
// scan: the candidate entries of rows[t], t < nrows (jacobi_rows.hpp's
//   RowScan: count, first candidate, the others appended to the shared list
//   of `cap` entries through *extra_n).
hipError_t launch_jacobi_rotate(double2* A, uint32_t d, uint32_t p, uint32_t q, double c, double sr, double si,
                                double npp, double nqq, double2* colp, double2* colq, uint32_t* extra_n, uint32_t ncol,
                                hipStream_t st);
// out[i] = Re A(i, i) for a d x d row-major double2 matrix.
hipError_t launch_diag_real(const double2* A, uint32_t d, double* out, hipStream_t st);
hipError_t launch_jacobi_scan(const double2* A, uint32_t d, const uint32_t* rows, uint32_t nrows, uint32_t* count,
                              uint32_t* first_j, double2* first_v, uint32_t* extra_n, uint32_t cap, uint32_t* extra_t,
                              uint32_t* extra_j, double2* extra_v, hipStream_t st);

}  // namespace iqg
