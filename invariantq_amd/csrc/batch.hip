// Batched variational sweep: many hardware-efficient-ansatz circuits of uneven
// size and depth, energy and full adjoint gradient of each, in ONE persistent
// launch.  No reference counterpart (the reference evaluates one circuit per
// simulator call, invariantq/benchmarks/vqe.py:44-62).
//
// Circuit c: n_c qubits, L_c layers of  RY(theta) on every qubit, RZ(phi) on
// every qubit, CZ on (q, q+1) and (n-1, 0); observable the open-chain TFIM
// H = sum_{q<n-1} Z_q Z_{q+1} + hx * sum_q X_q.  Parameters of layer l:
// theta_q = p[2nl + q], phi_q = p[2nl + n + q] (the layout of iq_circuit_*).
//
// One workgroup owns one circuit at a time: workgroup b starts with circuit b
// of the host's order (longest first), then takes the next from an atomic
// queue.  With b = the circuit's tile width:
//   n <= b  psi and lambda live in LDS for the whole circuit;
//   n >  b  they live in a per-workgroup HBM scratch, and a sweep is L + 1
//           tiled passes that alternate between kind A (the b low qubits,
//           contiguous) and kind B (the n-b high qubits plus the low 2b-n
//           ones, runs of 2^(2b-n) amplitudes).  Layer l's rotations go half
//           in pass l (the qubits of its kind) and half in pass l + 1, so pass
//           p applies layer p-1's second half, that layer's diagonal and layer
//           p's first half.  The reverse sweep runs the passes backwards.
// A layer is RY on every qubit (real rotations), then its diagonal D = CZ ring
// times the RZ set: one phase per amplitude, e^{i sum_q phi_q (g_q - 1/2)} with
// the CZ sign, applied in the pass that completes the layer's rotations
// (D^dagger in the reverse sweep).  Undoing D first leaves psi and lambda at
// the RY outputs, where dE/dtheta_q = Im<lambda|Y_q psi> and dE/dphi_q =
// Im<lambda|Z_q psi>; rotations on the other qubits, applied to both, leave a
// qubit's overlaps unchanged, so each half-layer's overlaps are taken where it
// is undone.  Inside a tile, rotations go in register groups of GW qubits, one
// LDS round trip per group.  Gradient inner products accumulate in FP64 per
// lane, reduce across the wave with shuffles into per-wave slots (one bank per
// layer parity), and are summed over waves in a fixed order once a layer is
// undone: deterministic, no atomics on results.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <type_traits>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"

namespace iqg {

namespace {

struct BatchCircuit {
    int32_t n;
    int32_t layers;
    int32_t tile_bits;
    int32_t pad;
    int64_t param_off;  // into params and grads (2 * n * layers values)
    int64_t out_index;  // energy slot
};

// Gradient slots per wave: two banks (layer parity) of 64, two per qubit
// (2q: dE/dphi_q, 2q + 1: dE/dtheta_q, n <= 30).  One tiled pass can undo the
// rotations of two layers, each into its own bank.
constexpr int kSlotBank = 64;
constexpr int kMaxSlots = 2 * kSlotBank;

// Per qubit and layer the kernel keeps the RY in scaled form.  With c, s the
// cosine and sine of half the angle, R0(u) = [[1, -u], [u, 1]], X the bit flip
// and Z = diag(1, -1):
//   mode 0 (|c| >= |s|): RY = c * R0(t),               t = s / c;
//   mode 1 (|s| >  |c|): RY = s * (-Z) X R0(-t),       t = c / s;
// so |t| <= 1 and R0 costs one fused multiply-add per output component (four
// per pair, where the plain rotation takes a multiply and a fused
// multiply-add each).  X is no arithmetic: a cube of amplitudes is stored with
// that bit of its positions flipped.  -Z is diagonal and commutes with every
// other qubit's rotation, so it joins the layer's diagonal phase.  The inverse
// is RY^dagger = f X^m R0(-t) (-Z)^m (m = mode, f = c or s): -Z first (with
// the inverse phase), then R0(-t), then the flip.  The factors f (|f| >= 1/sqrt 2)
// are not applied: every amplitude meets each qubit's f once, so a layer's
// rotations leave the whole state scaled by the product of its f, which the
// layer's phase restores.
// Entries per (layer, qubit): forward R0 parameter (t or -t), reverse (-t), f, m.
constexpr int kRotStride = 4;
// LDS phase tables: tile bits 0..6 and 7..12 (tiles of at most 2^13)
constexpr int kPhaseLo = 128, kPhaseHi = 64;

// (a, b) <- R0(u) (a, b) = (a - u b, u a + b)
template <typename R>
__device__ __forceinline__ void rot0(R u, cv<R>& a, cv<R>& b) {
    const cv<R> a0 = a;
    a.x = fma(-u, b.x, a0.x);
    a.y = fma(-u, b.y, a0.y);
    b.x = fma(u, a0.x, b.x);
    b.y = fma(u, a0.y, b.y);
}

// One block per circuit: rots[(param_off / 2 + n*l + q) * kRotStride + (0..3)]
// = (forward parameter, reverse parameter, f, mode) of the RY of qubit q in
// layer l, from FP64 angles.
template <typename R>
__global__ void k_prepare_rots(const BatchCircuit* __restrict__ circs, const double* __restrict__ params,
                               R* __restrict__ rots) {
    const BatchCircuit C = circs[blockIdx.x];
    const int n = C.n;
    for (int k = threadIdx.x; k < n * C.layers; k += blockDim.x) {
        const int l = k / n, q = k % n;
        const double th = params[C.param_off + 2 * n * l + q];
        const double c = cos(0.5 * th), s = sin(0.5 * th);
        const bool m1 = fabs(s) > fabs(c);
        const double t = m1 ? c / s : s / c;
        R* o = rots + (C.param_off / 2 + k) * kRotStride;
        o[0] = (R)(m1 ? -t : t);
        o[1] = (R)(-t);
        o[2] = (R)(m1 ? s : c);
        o[3] = m1 ? (R)1 : (R)0;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// global index of tile element j for pass kind (0 = whole state, 1 = A, 2 = B)
__device__ __forceinline__ uint64_t tile_global(int kind, uint64_t t, uint32_t j, int b, int n) {
    if (kind == 0) return j;
    if (kind == 1) return (t << b) | j;
    const int h = n - b, lo = b - h;
    const uint64_t jl = j & ((1u << lo) - 1), jh = j >> lo;
    return jl | (t << lo) | (jh << b);
}

__device__ __forceinline__ bool cz_ring_sign(uint64_t g, int n) {
    if (n < 2) return false;
    uint64_t m = g & (g >> 1) & ((1ULL << (n - 1)) - 1);
    int par = __popcll(m) & 1;
    if (n > 2) par ^= (int)(((g >> (n - 1)) & g) & 1ULL);
    return par != 0;
}

// GW: qubits per register group (the fusion width of the tile sweeps): 2..4
// for complex64, 2..3 for complex128.  tile_elems: amplitudes of LDS for the
// tiles (psi and lambda of the most demanding circuit of the launch).
template <typename R, int GW>
__global__ void __launch_bounds__(512) k_batch_hea(const BatchCircuit* __restrict__ circs, int ncirc,
                                                   const R* __restrict__ rots, const double* __restrict__ params,
                                                   double* __restrict__ grads,
                                                   double* __restrict__ energies, cv<R>* __restrict__ scratch,
                                                   uint64_t scratch_elems, int* __restrict__ queue, double hx,
                                                   uint32_t tile_elems) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = T >> 6;
    cv<R>* const tiles = reinterpret_cast<cv<R>*>(smem);
    double* wred = reinterpret_cast<double*>(tiles + tile_elems);  // [nw][kMaxSlots]
    cv<R>* ptab = reinterpret_cast<cv<R>*>(wred + nw * kMaxSlots);  // [kPhaseLo + kPhaseHi]
    __shared__ int s_next;
    cv<R>* gpsi = scratch + (uint64_t)blockIdx.x * 2 * scratch_elems;
    cv<R>* glam = gpsi + scratch_elems;

    // Workgroup b's first circuit is circuit b (the host sorted them longest
    // first), later ones come from the queue.  Were the first ones taken from
    // the queue too, which long circuits end up sharing a compute unit would
    // depend on the order the workgroups reach it (measured 180 against 211 ms
    // on one batch of 4096, run to run).
    bool first_round = true;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_next = first_round ? (int)blockIdx.x : (int)gridDim.x + atomicAdd(queue, 1);
        first_round = false;
        __syncthreads();
        // uniform: the circuit, its shape and its rotation table sit in SGPRs,
        // and the per-qubit rotation mode is a scalar branch
        const int ci = __builtin_amdgcn_readfirstlane(s_next);
        if (ci >= ncirc) break;
        const BatchCircuit C = circs[ci];
        const int n = C.n, L = C.layers;
        const int b = min(C.tile_bits, n);
        const bool resident = (n <= b);
        const uint32_t TS = 1u << b;
        const uint64_t N = 1ULL << n;
        const R* RT = rots + (C.param_off / 2) * kRotStride;  // [l][q][t, mode, f, f']
        double* G = grads + C.param_off;
        const uint64_t ntA = resident ? 1 : (N >> b);
        const int h = n - b;
        cv<R>* const tp = tiles;       // psi's tile
        cv<R>* const tl = tiles + TS;  // lambda's
        // Reverse sweep: psi and lambda in memory are the true ones divided by
        // the f' of the qubits already undone in the layer, so their overlaps
        // are off by the square of that product: sc_run, by which each group's
        // sums are multiplied as they are flushed.
        double sc_run = 1.0;

        // LDS position of tile element j: the low four index bits XOR the next
        // four, so sixteen lanes reading sixteen elements of a register group
        // hit distinct 8-byte slots whichever tile bits the group holds
        auto P = [](uint32_t j) __attribute__((always_inline)) -> uint32_t { return j ^ ((j >> 4) & 15u); };
        // qubits per register group, forward and reverse (sub-cubes of 2^GW
        // elements; the default is 4 for complex64 and 3 for complex128)
        constexpr int GM = GW;
        constexpr int GA = GW;
        // tile loads: U elements per thread per round (complex128: 8, so
        // nothing spills); with TS = U * T a round is one sub-cube over tile
        // bits RB .. RB + RU - 1 (the register group, when the group width is RU)
        constexpr int U = sizeof(R) == 4 ? 16 : 8, RU = sizeof(R) == 4 ? 4 : 3;
        const int RB = __builtin_ctz((uint32_t)T);
        const bool regq = !resident && TS == (uint32_t)U * (uint32_t)T && GW == RU;
        // The layer's diagonal D = CZ ring * prod_q RZ_q(phi_q): amplitude g takes
        // (-1)^cz(g) * e^{i sum_q phi_q (g_q - 1/2)}.  The phase of tile element j
        // (j < TS) is ptab[j & 127] * ptab[128 + (j >> 7)] times the CZ sign,
        // conjugated for D^dagger.
        auto phase_at = [&](int kind, uint64_t t, uint32_t j, bool dag, R& pr, R& pi) __attribute__((always_inline)) {
            const cv<R> a = ptab[j & 127u], c = ptab[kPhaseLo + (j >> 7)];
            pr = a.x * c.x - a.y * c.y;
            pi = a.x * c.y + a.y * c.x;
            if (cz_ring_sign(tile_global(kind, t, j, b, n), n)) {
                pr = -pr;
                pi = -pi;
            }
            if (dag) pi = -pi;
        };
        auto cmul = [](cv<R> v, R pr, R pi) __attribute__((always_inline)) {
            cv<R> o;
            o.x = v.x * pr - v.y * pi;
            o.y = v.x * pi + v.y * pr;
            return o;
        };
        // One sub-cube's reverse step: x (psi) and y (lambda) hold the 2^G
        // elements that differ in tile bits t0 .. t0 + G - 1 (qubit = tile bit
        // + qoff); the products go into az, aa, then RY^dagger on both.
        auto adj_cube = [&](auto Gc, auto& x, auto& y, int t0, int qoff, int l, double* az,
                            double* aa) __attribute__((always_inline)) {
            constexpr int G = decltype(Gc)::value;
            constexpr uint32_t D = 1u << G;
            R w[D], wall;
#pragma unroll
            for (uint32_t m = 0; m < D; ++m) {
                w[m] = fma(y[m].x, x[m].y, -(y[m].y * x[m].x));
                wall = m ? wall + w[m] : w[m];
            }
#pragma unroll
            for (int jb = G - 1; jb >= 0; --jb) {
                const uint32_t bit = 1u << jb;
                // one fused multiply-add per product term, chained through the sums
                R w1 = 0, pa = 0;
                bool first = true;
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) {
                    if (m & bit) continue;
                    const cv<R>&a0 = x[m], &a1 = x[m | bit], &b0 = y[m], &b1 = y[m | bit];
                    w1 = first ? w[m | bit] : w1 + w[m | bit];
                    pa = first ? b1.x * a0.x : fma(b1.x, a0.x, pa);
                    first = false;
                    pa = fma(b1.y, a0.y, pa);
                    pa = fma(-b0.x, a1.x, pa);
                    pa = fma(-b0.y, a1.y, pa);
                }
                az[jb] += (double)(wall - 2 * w1);
                aa[jb] += (double)pa;
            }
#pragma unroll
            for (int jb = G - 1; jb >= 0; --jb) {
                const R u = RT[(n * l + (t0 + jb + qoff)) * kRotStride + 1];
                const uint32_t bit = 1u << jb;
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) {
                    if (m & bit) continue;
                    rot0<R>(u, x[m], x[m | bit]);
                    rot0<R>(u, y[m], y[m | bit]);
                }
            }
        };
        // product of f^2 over the qubits of a group: the factor by which the
        // squared scale of psi and lambda grows once the group is undone
        auto group_scale = [&](int t0, int G, int qoff, int l) __attribute__((always_inline)) -> double {
            double f2 = 1.0;
            for (int jb = 0; jb < G; ++jb) {
                const double f = (double)RT[(n * l + (t0 + jb + qoff)) * kRotStride + 2];
                f2 *= f * f;
            }
            return f2;
        };
        // product of the layer's f
        auto layer_factor = [&](int l) __attribute__((always_inline)) -> double {
            double f = 1.0;
            for (int q = 0; q < n; ++q) f *= (double)RT[(n * l + q) * kRotStride + 2];
            return f;
        };
        // the group's mode-1 qubits as bits of the group's local index: the
        // flip X applied as the group's cube is stored
        auto group_flip = [&](int t0, int G, int qoff, int l) __attribute__((always_inline)) -> uint32_t {
            uint32_t jm = 0;
            for (int jb = 0; jb < G; ++jb)
                if (RT[(n * l + (t0 + jb + qoff)) * kRotStride + 3] != (R)0) jm |= 1u << jb;
            return jm;
        };
        // wave sums of a group's 2G products (qubit jb: z, a at 2 jb + 0, 1)
        // into the wave's slots, reduce-scatter: offsets 32, 16 and 8 halve the
        // values a lane carries (8 -> 4 -> 2 -> 1), offsets 4..1 finish the last
        auto flush_adj = [&](auto Gc, const double* az, const double* aa, int t0, int qoff, int l) __attribute__((always_inline)) {
            constexpr int G = decltype(Gc)::value;
            double acc[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int jb = k / 2;
                acc[k] = jb >= G ? 0.0 : (k % 2 == 0 ? az[jb < G ? jb : 0] : aa[jb < G ? jb : 0]);
            }
            const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0, h3 = (lane & 8) != 0;
            double v4[4], v2[2];
#pragma unroll
            for (int i = 0; i < 4; ++i) v4[i] = (h5 ? acc[4 + i] : acc[i]) + __shfl_xor(h5 ? acc[i] : acc[4 + i], 32, 64);
#pragma unroll
            for (int i = 0; i < 2; ++i) v2[i] = (h4 ? v4[2 + i] : v4[i]) + __shfl_xor(h4 ? v4[i] : v4[2 + i], 16, 64);
            double xs = (h3 ? v2[1] : v2[0]) + __shfl_xor(h3 ? v2[0] : v2[1], 8, 64);
#pragma unroll
            for (int off = 4; off > 0; off >>= 1) xs += __shfl_xor(xs, off, 64);
            const int k = (h5 ? 4 : 0) + (h4 ? 2 : 0) + (h3 ? 1 : 0);
            if ((lane & 7) == 0 && k < 2 * G) wred[wave * kMaxSlots + (l & 1) * kSlotBank + 2 * (t0 + qoff) + k] += xs * sc_run;
            sc_run *= group_scale(t0, G, qoff, l);
        };
        // --- tile transfer helpers: eight or sixteen elements per thread in
        // flight before any is written on, so a tile costs about one memory
        // latency, not one per element
        // psi (and lambda) tile in one round of loads: up to 16 elements per
        // thread per array in flight (64 KiB per CU for a 2^13 complex64 tile;
        // 8 for complex128)
        // dg: 0 as stored, 1 with D, 2 with D^dagger on the way into LDS.
        // rop (register group, when TS = U * T: each thread's elements j0 + u T
        // are one sub-cube over tile bits RB .. RB + RU - 1, RB = log2 T):
        // 1 applies those qubits' RY (forward), 2 their reverse step (psi and
        // lambda), in registers before the tile goes to LDS; qubit = tile bit + qoff.
        auto load2 = [&](int kind, uint64_t t, bool two, int dg, int rop, int qoff, int l) __attribute__((always_inline)) {
            for (uint32_t j0 = tid; j0 < TS; j0 += U * T) {
                cv<R> v[U], w[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {  // every element assigned: the arrays stay in registers
                    const uint64_t gi = tile_global(kind, t, min(j0 + u * T, TS - 1), b, n);
                    v[u] = gpsi[gi];
                    w[u] = two ? glam[gi] : v[u];
                }
                if (dg) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const uint32_t j = j0 + u * T;
                        if (j < TS) {
                            R pr, pi;
                            phase_at(kind, t, j, dg == 2, pr, pi);
                            v[u] = cmul(v[u], pr, pi);
                            if (two) w[u] = cmul(w[u], pr, pi);
                        }
                    }
                }
                // the register group's flips: element u goes to u ^ jm
                const uint32_t pjm = rop ? P(group_flip(RB, RU, qoff, l) << RB) : 0u;
                if (rop == 1) {
#pragma unroll
                    for (int i = 0; i < RU; ++i) {
                        const R u0 = RT[(n * l + (RB + i + qoff)) * kRotStride];
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (!(u & (1 << i))) rot0<R>(u0, v[u], v[u | (1 << i)]);
                    }
                } else if (rop == 2) {
                    double az[RU], aa[RU];
#pragma unroll
                    for (int i = 0; i < RU; ++i) az[i] = aa[i] = 0.0;
                    adj_cube(std::integral_constant<int, RU>{}, v, w, RB, qoff, l, az, aa);
                    flush_adj(std::integral_constant<int, RU>{}, az, aa, RB, qoff, l);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t j = j0 + u * T;
                    if (j < TS) {
                        tp[P(j) ^ pjm] = v[u];
                        if (two) tl[P(j) ^ pjm] = w[u];
                    }
                }
            }
        };
        auto load = [&](int kind, uint64_t t, cv<R>* dst, const cv<R>* src) __attribute__((always_inline)) {
            for (uint32_t j0 = tid; j0 < TS; j0 += 8 * T) {
                cv<R> v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {  // every v[u] assigned: the array stays in registers
                    const uint32_t j = min(j0 + u * T, TS - 1);
                    v[u] = src[tile_global(kind, t, j, b, n)];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint32_t j = j0 + u * T;
                    if (j < TS) dst[P(j)] = v[u];
                }
            }
        };
        // dg as load2, on the way out of LDS
        auto store = [&](int kind, uint64_t t, const cv<R>* src, cv<R>* dst, int dg) __attribute__((always_inline)) {
            for (uint32_t j0 = tid; j0 < TS; j0 += 8 * T) {
                cv<R> v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint32_t j = min(j0 + u * T, TS - 1);
                    v[u] = src[P(j)];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint32_t j = j0 + u * T;
                    if (j < TS) {
                        if (dg) {
                            R pr, pi;
                            phase_at(kind, t, j, dg == 2, pr, pi);
                            v[u] = cmul(v[u], pr, pi);
                        }
                        dst[tile_global(kind, t, j, b, n)] = v[u];
                    }
                }
            }
        };
        // Rotations of tile bits [t0, t0 + G) (qubit = tile bit + qoff): one LDS
        // round trip per group, not per qubit.  Each thread takes sub-cubes of
        // 2^G elements that differ only in those bits into registers and
        // applies the G rotations there.  G is a compile-time constant so the
        // sub-cube stays in registers.
        auto rot_g = [&](auto Gc, int t0, int qoff, int l) __attribute__((always_inline)) {
            constexpr int G = decltype(Gc)::value;
            constexpr uint32_t D = 1u << G;
            const uint32_t cubes = TS >> G, lo = (1u << t0) - 1;
            const uint32_t pj = P(group_flip(t0, G, qoff, l) << t0);
            for (uint32_t s = tid; s < cubes; s += T) {
                const uint32_t base = ((s & ~lo) << G) | (s & lo);
                const uint32_t pb0 = P(base);  // P is linear over disjoint bits
                cv<R> x[D];
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) x[m] = tp[pb0 ^ P(m << t0)];
#pragma unroll
                for (int jb = 0; jb < G; ++jb) {
                    const R u0 = RT[(n * l + (t0 + jb + qoff)) * kRotStride];
#pragma unroll
                    for (uint32_t m = 0; m < D; ++m)
                        if (!(m & (1u << jb))) rot0<R>(u0, x[m], x[m | (1u << jb)]);
                }
                const uint32_t pbs = pb0 ^ pj;  // stored with the mode-1 bits flipped
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) tp[pbs ^ P(m << t0)] = x[m];
            }
            __syncthreads();
        };
        // Tile bits [first, end) cut into groups of at most gm, of about equal
        // size (a lone one-qubit group costs nearly a full LDS round trip)
        auto for_groups = [&](int first, int end, int gm, auto&& f) __attribute__((always_inline)) {
            const int len = end - first, ng = (len + gm - 1) / gm;
            int t0 = first;
            for (int k = 0; k < ng; ++k) {
                const int g = len / ng + (k < len % ng ? 1 : 0);
                f(t0, g);
                t0 += g;
            }
        };
        // dispatch on the group width; callers pass 1 <= g <= GM
        auto rot_group = [&](int t0, int g, int qoff, int l) __attribute__((always_inline)) {
            static_assert(GM >= 2 && GM <= 4, "group width");
            switch (g) {
                case 1: rot_g(std::integral_constant<int, 1>{}, t0, qoff, l); break;
                case 2: rot_g(std::integral_constant<int, 2>{}, t0, qoff, l); break;
                case 3: if constexpr (GM >= 3) rot_g(std::integral_constant<int, 3>{}, t0, qoff, l); break;
                case 4: if constexpr (GM >= 4) rot_g(std::integral_constant<int, 4>{}, t0, qoff, l); break;
                default: break;
            }
        };
        // Reverse steps of tile bits [t0, t0 + G) on psi and lambda.  The layer's
        // CZ ring and RZ set are already undone (one phase per amplitude), so
        // psi and lambda sit at the RY outputs, where for qubit q
        //   dE/dtheta = Im<lambda|Y psi> = sum Re(l1* p0) - Re(l0* p1),
        //   dE/dphi   = Im<lambda|Z psi> (Z commutes with the RZ set)
        //             = sum over bit = 0 minus bit = 1 of w_m = Im(conj(l_m) p_m);
        // then RY^dagger (real) on psi and lambda.  Every qubit's products come
        // from the same registers (rotations on the group's other qubits,
        // applied to both, leave them unchanged), summed in the state's
        // precision over a sub-cube, then in FP64.
        auto adj_g = [&](auto Gc, int t0, int qoff, int l) __attribute__((always_inline)) {
            constexpr int G = decltype(Gc)::value;
            constexpr uint32_t D = 1u << G;
            const uint32_t cubes = TS >> G, lo = (1u << t0) - 1;
            double az[G], aa[G];
#pragma unroll
            for (int jb = 0; jb < G; ++jb) az[jb] = aa[jb] = 0.0;
            const uint32_t pj = P(group_flip(t0, G, qoff, l) << t0);
            for (uint32_t s = tid; s < cubes; s += T) {
                const uint32_t base = ((s & ~lo) << G) | (s & lo);
                const uint32_t pb0 = P(base);  // P is linear over disjoint bits
                cv<R> x[D], y[D];
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) {
                    x[m] = tp[pb0 ^ P(m << t0)];
                    y[m] = tl[pb0 ^ P(m << t0)];
                }
                adj_cube(Gc, x, y, t0, qoff, l, az, aa);
                const uint32_t pbs = pb0 ^ pj;  // stored with the mode-1 bits flipped
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) {
                    tp[pbs ^ P(m << t0)] = x[m];
                    tl[pbs ^ P(m << t0)] = y[m];
                }
            }
            flush_adj(Gc, az, aa, t0, qoff, l);
            __syncthreads();
        };
        // dispatch on the group width; callers pass 1 <= g <= GA
        auto adj_group = [&](int t0, int g, int qoff, int l) __attribute__((always_inline)) {
            static_assert(GA >= 2 && GA <= 4, "group width");
            switch (g) {
                case 1: adj_g(std::integral_constant<int, 1>{}, t0, qoff, l); break;
                case 2: adj_g(std::integral_constant<int, 2>{}, t0, qoff, l); break;
                case 3: if constexpr (GA >= 3) adj_g(std::integral_constant<int, 3>{}, t0, qoff, l); break;
                case 4: if constexpr (GA >= 4) adj_g(std::integral_constant<int, 4>{}, t0, qoff, l); break;
                default: break;
            }
        };
        // One pass over the tiles of a decomposition (kind 1: the b low qubits,
        // 2: the high qubits): each tile goes global -> LDS, through `body`,
        // and back; pre(t) writes tile t's LDS tables, and dl, ds say which
        // diagonal the loads and the stores take (as load2).  (Prefetching the
        // next tile into registers across `body` needs 64 more VGPRs per lane
        // and spilled at 512 threads.)
        auto tile_pass = [&](int kind, bool two, int dl, int ds, int rop, int qoff, int l, auto&& pre,
                             auto&& body) __attribute__((always_inline)) {
            for (uint64_t t = 0; t < ntA; ++t) {
                pre(t);  // the last tile's users passed the barrier below
                load2(kind, t, two, dl, rop, qoff, l);
                __syncthreads();
                body(t);
                store(kind, t, tp, gpsi, ds);
                if (two) store(kind, t, tl, glam, ds);
                __syncthreads();
            }
        };
        // The tables of phase_at, made in FP64 (angle sums, then sincos) for the
        // tile's bit-to-qubit map, the second with the angles of the qubits the
        // tile number fixes folded in.
        auto build_phase = [&](int kind, uint64_t t, int l, double scale) __attribute__((always_inline)) {
            const double* ph = params + C.param_off + 2 * n * l + n;
            // kind 2: tile bits [0, lo) on qubits [0, lo), the rest on [b, n);
            // kind 1: tile bit k on qubit k; kind 0: the whole state
            const int lo = 2 * b - n;
            const int nb = kind == 0 ? n : b;
            auto qubit = [&](int k) { return kind == 2 && k >= lo ? k - lo + b : k; };
            const uint32_t na = TS < 128u ? TS : 128u, nbt = TS > 128u ? TS >> 7 : 1u;
            for (uint32_t i = tid; i < na + nbt; i += T) {
                // qubit q at bit value g: phi_q (g - 1/2), and pi where a mode-1
                // rotation's -Z meets g = 0
                auto term = [&](int q, uint32_t g) {
                    const bool m1 = RT[(n * l + q) * kRotStride + 3] != (R)0;
                    return ph[q] * ((double)g - 0.5) + (m1 && !g ? M_PI : 0.0);
                };
                double ang = 0.0;
                if (i < na) {
                    for (int k = 0; k < 7 && k < nb; ++k) ang += term(qubit(k), (i >> k) & 1u);
                } else {
                    const uint32_t e = i - na;
                    for (int k = 7; k < nb; ++k) ang += term(qubit(k), (e >> (k - 7)) & 1u);
                    if (kind == 2)  // tile number bit k on qubit lo + k
                        for (int k = 0; k < n - b; ++k) ang += term(lo + k, (uint32_t)(t >> k) & 1u);
                    if (kind == 1)  // tile number bit k on qubit b + k
                        for (int k = 0; k < n - b; ++k) ang += term(b + k, (uint32_t)(t >> k) & 1u);
                }
                double sv, cvl;
                sincos(ang, &sv, &cvl);
                if (i < na) {  // the state's pending scale, once per phase
                    sv *= scale;
                    cvl *= scale;
                }
                cv<R> v;
                v.x = (R)cvl;
                v.y = (R)sv;
                ptab[i < na ? i : kPhaseLo + (i - na)] = v;
            }
        };
        // D (forward) or D^dagger (reverse, on psi and lambda) over the tile
        // in LDS (kind 0: the resident state)
        auto diag = [&](int kind, uint64_t t, bool both) __attribute__((always_inline)) {
            for (uint32_t j = tid; j < TS; j += T) {
                R pr, pi;
                phase_at(kind, t, j, both, pr, pi);
                tp[P(j)] = cmul(tp[P(j)], pr, pi);
                if (both) tl[P(j)] = cmul(tl[P(j)], pr, pi);
            }
            __syncthreads();
        };
        auto clear_slots = [&]() __attribute__((always_inline)) {
            for (int i = tid; i < nw * kMaxSlots; i += T) wred[i] = 0.0;
            __syncthreads();
        };
        // layer gradients from the layer's bank of slots (2q: dE/dphi_q,
        // 2q+1: dE/dtheta_q), summed over waves in a fixed order; the bank is
        // cleared for layer l - 2
        auto flush_slots = [&](int l) __attribute__((always_inline)) {
            const int bank = (l & 1) * kSlotBank;
            for (int q = tid; q < n; q += T) {
                double z = 0.0, va = 0.0;
                for (int w = 0; w < nw; ++w) {
                    z += wred[w * kMaxSlots + bank + 2 * q];
                    va += wred[w * kMaxSlots + bank + 2 * q + 1];
                }
                G[2 * n * l + n + q] = z;
                // a mode-1 rotation's -Z, applied before the overlaps, turns Y into -Y
                G[2 * n * l + q] = RT[(n * l + q) * kRotStride + 3] != (R)0 ? -va : va;
            }
            __syncthreads();
            for (int i = tid; i < nw * kSlotBank; i += T) wred[(i / kSlotBank) * kMaxSlots + bank + i % kSlotBank] = 0.0;
            __syncthreads();
        };

        // --- forward: psi = U|0>
        if (resident) {
            for (uint32_t j = tid; j < TS; j += T) {
                cv<R> v;
                v.x = (j == 0) ? (R)1 : (R)0;
                v.y = 0;
                tp[P(j)] = v;
            }
            __syncthreads();
        } else {
            for (uint64_t j = tid; j < N; j += T) {
                cv<R> v;
                v.x = (j == 0) ? (R)1 : (R)0;
                v.y = 0;
                gpsi[j] = v;
            }
            __syncthreads();
        }
        if (resident) {
            for (int l = 0; l < L; ++l) {
                // the layer's rotations leave the state divided by ff, restored
                // with the diagonal's phase
                const double ff = layer_factor(l);
                for_groups(0, n, GM, [&](int t0, int g) { rot_group(t0, g, 0, l); });
                build_phase(0, 0, l, ff);
                __syncthreads();
                diag(0, 0, false);
            }
        } else {
            // L + 1 tiled passes: pass p tiles kind A (the b low qubits) for even
            // p, kind B (the high qubits) for odd p.  Layer l's RY go half in
            // pass l (the qubits of that pass's kind) and half in pass l + 1,
            // whose kind holds the other qubits; rotations on different qubits
            // commute, so pass p applies layer p - 1's second half, then D_{p-1}
            // (every qubit of layer p - 1 rotated), then layer p's first half:
            // each pass serves two layers, where one pass per half-layer would
            // move the state 2L times.
            for (int p = 0; p <= L; ++p) {
                const int kind = (p & 1) ? 2 : 1, qo = kind == 2 ? h : 0, g0 = kind == 2 ? b - h : 0;
                const bool rg = regq && (kind == 1 || b - h <= RB);
                const bool prev = p >= 1, next = p < L;
                const int lr = prev ? p - 1 : 0;  // the layer whose rotations the pass starts with
                const double ff = prev ? layer_factor(p - 1) : 1.0;
                // the register group's rotations as the tile is loaded; D_{L-1}
                // (the last pass) as it is stored
                tile_pass(kind, false, 0, prev && !next ? 1 : 0, rg ? 1 : 0, qo, lr, [&](uint64_t t) {
                    if (prev) build_phase(kind, t, p - 1, ff);  // the loads' barrier orders it first
                }, [&](uint64_t t) {
                    for_groups(g0, rg ? RB : b, GM, [&](int t0, int g) { rot_group(t0, g, qo, lr); });
                    if (prev && next) {
                        diag(kind, t, false);
                        for_groups(g0, b, GM, [&](int t0, int g) { rot_group(t0, g, qo, p); });
                    }
                });
            }
        }

        // --- lambda = H psi, energy = Re <psi|lambda>
        double e = 0.0;
        auto hdiag = [&](uint64_t g) __attribute__((always_inline)) {
            const uint64_t diff = (g ^ (g >> 1)) & ((1ULL << (n - 1)) - 1);
            return (double)(n - 1) - 2.0 * (double)__popcll(diff);
        };
        if (resident) {
            for (uint32_t j = tid; j < TS; j += T) {
                const cv<R> v = tp[P(j)];
                double sr = hdiag(j) * (double)v.x, si = hdiag(j) * (double)v.y;
                for (int q = 0; q < n; ++q) {
                    const cv<R> w = tp[P(j ^ (1u << q))];
                    sr += hx * (double)w.x;
                    si += hx * (double)w.y;
                }
                cv<R> o;
                o.x = (R)sr;
                o.y = (R)si;
                tl[P(j)] = o;
                e += (double)v.x * sr + (double)v.y * si;
            }
            __syncthreads();
        } else {
            for (uint64_t t = 0; t < ntA; ++t) {  // diagonal + X on the b low qubits
                load(1, t, tp, gpsi);
                __syncthreads();
                for (uint32_t j = tid; j < TS; j += T) {
                    const uint64_t g = tile_global(1, t, j, b, n);
                    const cv<R> v = tp[P(j)];
                    double sr = hdiag(g) * (double)v.x, si = hdiag(g) * (double)v.y;
                    for (int q = 0; q < b; ++q) {
                        const cv<R> w = tp[P(j ^ (1u << q))];
                        sr += hx * (double)w.x;
                        si += hx * (double)w.y;
                    }
                    cv<R> o;
                    o.x = (R)sr;
                    o.y = (R)si;
                    glam[g] = o;
                }
                __syncthreads();
            }
            for (uint64_t t = 0; t < ntA; ++t) {  // X on the high qubits, then the energy
                load(2, t, tp, gpsi);
                load(2, t, tl, glam);
                __syncthreads();
                for (uint32_t j = tid; j < TS; j += T) {
                    const cv<R> v = tp[P(j)];
                    cv<R> o = tl[P(j)];
                    double sr = (double)o.x, si = (double)o.y;
                    for (int q = b; q < n; ++q) {
                        const cv<R> w = tp[P(j ^ (1u << (q - b + (b - h))))];
                        sr += hx * (double)w.x;
                        si += hx * (double)w.y;
                    }
                    o.x = (R)sr;
                    o.y = (R)si;
                    tl[P(j)] = o;
                    e += (double)v.x * sr + (double)v.y * si;
                }
                __syncthreads();
                store(2, t, tl, glam, 0);
                __syncthreads();
            }
        }
        e = wave_sum(e);
        if (lane == 0) wred[wave * kMaxSlots] = e;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += wred[w * kMaxSlots];
            energies[C.out_index] = s;
        }
        __syncthreads();

        // --- reverse sweep
        clear_slots();
        if (resident) {
            for (int l = L - 1; l >= 0; --l) {
                // psi and lambda come divided by the later layer's product of
                // f', restored with D^dagger's phase
                const double fpend = l + 1 < L ? layer_factor(l + 1) : 1.0;
                build_phase(0, 0, l, fpend);
                __syncthreads();
                diag(0, 0, true);
                sc_run = 1.0;
                for_groups(0, n, GA, [&](int t0, int g) { adj_group(t0, g, 0, l); });
                flush_slots(l);
            }
        } else {
            // the forward passes in reverse: pass p undoes layer p's first half
            // (D_p^dagger was applied by pass p + 1), then D_{p-1}^dagger, then
            // layer p - 1's second half.  Layer p's overlaps carry the square of
            // the f' of its qubits already undone: sc_run continues from the end
            // of pass p + 1, where its second half was undone; every tile runs
            // the same sequence, so it restarts at the pass's value per tile.
            double sc_carry = 1.0;
            for (int p = L; p >= 0; --p) {
                const int kind = (p & 1) ? 2 : 1, qo = kind == 2 ? h : 0, g0 = kind == 2 ? b - h : 0;
                const bool rg = regq && (kind == 1 || b - h <= RB);
                const bool prev = p >= 1, next = p < L;
                const int lr = next ? p : p - 1;  // the layer whose rotations the pass undoes first
                // psi and lambda come divided by layer p's product of f' (every
                // qubit of it undone), restored with D_{p-1}^dagger's phase
                const double fpend = next ? layer_factor(p) : 1.0;
                const double sc_in = next ? sc_carry : 1.0;
                // the last pass's D_{L-1}^dagger as the tile is loaded
                tile_pass(kind, true, next ? 0 : 2, 0, rg ? 2 : 0, qo, lr, [&](uint64_t t) {
                    if (prev) build_phase(kind, t, p - 1, fpend);
                    if (!next) __syncthreads();  // the loads use the table
                    sc_run = sc_in;
                }, [&](uint64_t t) {
                    for_groups(g0, rg ? RB : b, GA, [&](int t0, int g) { adj_group(t0, g, qo, lr); });
                    if (prev && next) {
                        diag(kind, t, true);
                        sc_run = 1.0;
                        for_groups(g0, b, GA, [&](int t0, int g) { adj_group(t0, g, qo, p - 1); });
                    }
                });
                sc_carry = sc_run;
                if (next) flush_slots(p);  // both halves of layer p undone
            }
        }
    }
}

__global__ void k_zero_int(int* p) { *p = 0; }

}  // namespace

struct BatchContext {
    int prec = kC64;
    int dev = 0;
    hipStream_t stream = nullptr;
    int* queue = nullptr;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    BatchCircuit* circs = nullptr;
    size_t circ_cap = 0;
    double* params = nullptr;
    double* grads = nullptr;
    size_t param_cap = 0;
    double* energies = nullptr;
    size_t energy_cap = 0;
    void* rots = nullptr;
    size_t rots_cap = 0;
    int group_width = 0;  // qubits per register group, 0 = the precision's default
    ~BatchContext() {
        if (rots) hipFree(rots);
        if (stream) hipStreamSynchronize(stream);
        if (queue) hipFree(queue);
        if (scratch) hipFree(scratch);
        if (circs) hipFree(circs);
        if (params) hipFree(params);
        if (grads) hipFree(grads);
        if (energies) hipFree(energies);
        if (stream) hipStreamDestroy(stream);
    }
};

static bool grow(void** p, size_t& cap, size_t need) {
    if (cap >= need) return true;
    if (*p) hipFree(*p);
    *p = nullptr;
    cap = 0;
    if (hipMalloc(p, std::max<size_t>(need, 64)) != hipSuccess) {
        set_error(kErrOutOfMemory, "cannot allocate %zu bytes for the batch", need);
        return false;
    }
    cap = need;
    return true;
}

}  // namespace iqg

using namespace iqg;

extern "C" {

void* iq_batch_create(int precision) {
    if (precision != kC128 && precision != kC64) {
        set_error(kErrInvalidArgument, "unknown precision %d", precision);
        return nullptr;
    }
    BatchContext* b = new (std::nothrow) BatchContext();
    if (!b) return nullptr;
    b->prec = precision;
    hipGetDevice(&b->dev);
    if (!hip_ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream") ||
        !hip_ok(hipMalloc(&b->queue, sizeof(int)), "queue")) {
        delete b;
        return nullptr;
    }
    return b;
}

void iq_batch_free(void* bp) {
    if (!bp) return;
    DeviceGuard g(static_cast<BatchContext*>(bp)->dev);
    delete static_cast<BatchContext*>(bp);
}

// Largest tile width (qubits held in LDS) for the batch kernel in this precision.
int iq_batch_max_tile_bits(int precision) { return precision == kC128 ? 12 : 13; }

// Largest register-group width (qubits fused per register pass of the tile sweeps).
int iq_batch_max_group_width(int precision) { return precision == kC128 ? 3 : 4; }

int iq_batch_set_group_width(void* bp, int width) {
    if (!bp) return -1;
    BatchContext& B = *static_cast<BatchContext*>(bp);
    if (width != 0 && (width < 2 || width > iq_batch_max_group_width(B.prec))) {
        set_error(kErrInvalidArgument, "group width %d outside 2..%d (0 = default)", width,
                  iq_batch_max_group_width(B.prec));
        return -1;
    }
    B.group_width = width;
    return 0;
}

// Energies and adjoint gradients of ncirc ansatz circuits in one persistent launch.
//   nq[c], layers[c]: circuit shape (1 <= nq <= 30, layers >= 1);
//   params: concatenated, 2 * nq[c] * layers[c] values per circuit (in c order);
//   tile_bits[c]: LDS tile width for circuit c (0 = the largest), nq <= width keeps it resident;
//   block_threads: 256, 512 (default when 0) or 1024 ... limited to 512 by the kernel bounds;
//   grads_out: same layout as params; energies_out: ncirc values.
int iq_batch_hea_tfim(void* bp, size_t ncirc, const int32_t* nq, const int32_t* layers, const double* params, double hx,
                      const int32_t* tile_bits, int block_threads, int num_blocks, double* energies_out,
                      double* grads_out, double* elapsed_ms) {
    if (!bp || !nq || !layers || !params || !energies_out || !grads_out) return -1;
    BatchContext& B = *static_cast<BatchContext*>(bp);
    DeviceGuard dg(B.dev);
    if (ncirc == 0) return 0;
    const int bmax = iq_batch_max_tile_bits(B.prec);
    const int T = block_threads ? block_threads : 512;
    if (T != 128 && T != 256 && T != 512) {
        set_error(kErrInvalidArgument, "block_threads must be 128, 256 or 512");
        return -1;
    }
    std::vector<BatchCircuit> cs(ncirc);
    std::vector<int64_t> off(ncirc);
    int64_t total = 0;
    int nmax = 0;
    // LDS amplitudes the most demanding circuit needs: its psi and lambda
    // (resident, or tiles of its width)
    uint32_t tile_elems = 0;
    for (size_t c = 0; c < ncirc; ++c) {
        if (nq[c] < 2 || nq[c] > 30 || layers[c] < 1 || 3 * nq[c] > kMaxSlots) {
            set_error(kErrInvalidArgument, "circuit %zu: unsupported shape (n=%d, layers=%d)", c, nq[c], layers[c]);
            return -1;
        }
        int tb = (tile_bits && tile_bits[c]) ? tile_bits[c] : bmax;
        tb = std::min(tb, bmax);
        if (nq[c] > tb && 2 * tb < nq[c]) {
            set_error(kErrInvalidArgument, "circuit %zu: tile width %d too small for %d qubits (need 2*width >= n)",
                      c, tb, nq[c]);
            return -1;
        }
        if (tb < 6) tb = 6;
        tile_elems = std::max<uint32_t>(tile_elems, 2u << std::min(tb, (int)nq[c]));
        cs[c] = BatchCircuit{nq[c], layers[c], tb, 0, total, (int64_t)c};
        off[c] = total;
        total += 2LL * nq[c] * layers[c];
        if (nq[c] > tb) nmax = std::max(nmax, nq[c]);
    }
    // longest first: cost ~ layers * 2^n
    std::vector<size_t> order(ncirc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b2) {
        return (double)cs[a].layers * std::ldexp(1.0, cs[a].n) > (double)cs[b2].layers * std::ldexp(1.0, cs[b2].n);
    });
    std::vector<BatchCircuit> sorted(ncirc);
    for (size_t i = 0; i < ncirc; ++i) sorted[i] = cs[order[i]];
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const size_t elem = B.prec == kC128 ? 16 : 8;
    // LDS for the widest tile in use (psi and lambda) plus the per-wave
    // gradient slots; as many workgroups as fit a CU at that size run at once
    const size_t lds = (size_t)tile_elems * elem + (size_t)(T / 64) * kMaxSlots * sizeof(double) +
                       (size_t)(kPhaseLo + kPhaseHi) * elem;
    const int gw = B.group_width ? B.group_width : iq_batch_max_group_width(B.prec);
    const void* kern = B.prec == kC128 ? (gw == 2 ? (const void*)k_batch_hea<double, 2> : (const void*)k_batch_hea<double, 3>)
                       : gw == 2       ? (const void*)k_batch_hea<float, 2>
                       : gw == 3       ? (const void*)k_batch_hea<float, 3>
                                       : (const void*)k_batch_hea<float, 4>;
    hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    // Workgroups per compute unit, from the kernel's VGPRs and LDS and gfx950's
    // fixed resources (512 VGPRs per SIMD lane, 4 SIMDs, up to 8 waves per
    // SIMD).  Not hipOccupancyMaxActiveBlocksPerMultiprocessor: with another
    // HIP runtime in the process (PyTorch's) the device reports half the
    // register file and the query returns one workgroup where two fit.
    int per_cu = 1;
    {
        hipFuncAttributes fa{};
        int smem_cu = 0;
        hipDeviceGetAttribute(&smem_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
        if (hipFuncGetAttributes(&fa, kern) == hipSuccess && fa.numRegs > 0 && smem_cu > 0) {
            const int vgprs = (fa.numRegs + 7) / 8 * 8, waves_simd = (T / 64 + 3) / 4;
            const int by_regs = std::min(512 / vgprs, 8) / waves_simd, by_lds = smem_cu / (int)lds;
            per_cu = std::max(1, std::min(by_regs, by_lds));
        }
    }
    const int nblk = num_blocks > 0 ? num_blocks : std::max(1, cus * per_cu);
    const uint64_t scratch_elems = nmax ? (1ULL << nmax) : 1;
    if (!grow(reinterpret_cast<void**>(&B.circs), B.circ_cap, ncirc * sizeof(BatchCircuit)) ||
        !grow(reinterpret_cast<void**>(&B.params), B.param_cap, (size_t)total * 8) ||
        !grow(reinterpret_cast<void**>(&B.energies), B.energy_cap, ncirc * 8) ||
        !grow(&B.scratch, B.scratch_bytes, (size_t)nblk * 2 * scratch_elems * elem) ||
        !grow(&B.rots, B.rots_cap, (size_t)(total / 2) * kRotStride * (B.prec == kC128 ? 8 : 4)))
        return -1;
    if (B.grads) hipFree(B.grads);
    B.grads = nullptr;
    if (!hip_ok(hipMalloc(&B.grads, std::max<size_t>((size_t)total * 8, 64)), "grads")) return -1;
    hipStream_t st = B.stream;
    if (!hip_ok(hipMemcpyAsync(B.circs, sorted.data(), ncirc * sizeof(BatchCircuit), hipMemcpyHostToDevice, st), "H2D") ||
        !hip_ok(hipMemcpyAsync(B.params, params, (size_t)total * 8, hipMemcpyHostToDevice, st), "H2D"))
        return -1;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, B.queue);
    if (B.prec == kC128)
        hipLaunchKernelGGL(k_prepare_rots<double>, dim3((unsigned)ncirc), dim3(256), 0, st, B.circs, B.params,
                           (double*)B.rots);
    else
        hipLaunchKernelGGL(k_prepare_rots<float>, dim3((unsigned)ncirc), dim3(256), 0, st, B.circs, B.params,
                           (float*)B.rots);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, st);
#define IQ_BATCH_LAUNCH(RT, GWV, CV)                                                                              \
    hipLaunchKernelGGL((k_batch_hea<RT, GWV>), dim3(nblk), dim3(T), lds, st, B.circs, (int)ncirc, (const RT*)B.rots, \
                       B.params, B.grads, B.energies, (CV*)B.scratch, scratch_elems, B.queue, hx, tile_elems)
    if (B.prec == kC128) {
        if (gw == 2) IQ_BATCH_LAUNCH(double, 2, double2);
        else IQ_BATCH_LAUNCH(double, 3, double2);
    } else {
        if (gw == 2) IQ_BATCH_LAUNCH(float, 2, float2);
        else if (gw == 3) IQ_BATCH_LAUNCH(float, 3, float2);
        else IQ_BATCH_LAUNCH(float, 4, float2);
    }
#undef IQ_BATCH_LAUNCH
    if (!hip_ok(hipGetLastError(), "batch launch")) return -1;
    hipEventRecord(e1, st);
    std::vector<double> en(ncirc);
    if (!hip_ok(hipMemcpyAsync(grads_out, B.grads, (size_t)total * 8, hipMemcpyDeviceToHost, st), "D2H") ||
        !hip_ok(hipMemcpyAsync(en.data(), B.energies, ncirc * 8, hipMemcpyDeviceToHost, st), "D2H") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return -1;
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    for (size_t c = 0; c < ncirc; ++c) energies_out[c] = en[c];
    (void)off;
    return 0;
}

}  // extern "C"
