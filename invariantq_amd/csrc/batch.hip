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
// One workgroup owns one circuit at a time, taken from an atomic queue in the
// order the host sorted (longest first).  With b = the circuit's tile width:
//   n <= b  psi and lambda live in LDS for the whole circuit;
//   n >  b  they live in a per-workgroup HBM scratch and every ansatz layer is
//           two tiled passes: pass A tiles the b low qubits (contiguous), pass B
//           tiles the n-b high qubits plus the low 2b-n ones (runs of 2^(2b-n)
//           amplitudes) and also applies the CZ ring, which depends only on the
//           global index.  The reverse sweep runs B then A per layer.
// Inside a tile each qubit is one LDS sweep over 2^(b-1) pairs.  Gradient inner
// products Im<lambda|G psi> accumulate in FP64 per lane, reduce across the
// wave with shuffles into per-wave slots, and are summed over waves in a fixed
// order at the end of each layer: deterministic, no atomics on results.

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

constexpr int kMaxSlots = 96;  // gradient slots per layer: 3 per qubit (n <= 30)

template <typename R>
struct Rot {
    R u[8];  // re00 im00 re01 im01 re10 im10 re11 im11
};

template <typename R>
__device__ __forceinline__ void mul2(const R* u, R& x0r, R& x0i, R& x1r, R& x1i) {
    const R y0r = u[0] * x0r - u[1] * x0i + u[2] * x1r - u[3] * x1i;
    const R y0i = u[0] * x0i + u[1] * x0r + u[2] * x1i + u[3] * x1r;
    const R y1r = u[4] * x0r - u[5] * x0i + u[6] * x1r - u[7] * x1i;
    const R y1i = u[4] * x0i + u[5] * x0r + u[6] * x1i + u[7] * x1r;
    x0r = y0r; x0i = y0i; x1r = y1r; x1i = y1i;
}

// RZ(phi) RY(theta) and its inverse (RZ RY)^dagger, from FP64 angles.
template <typename R>
__device__ __forceinline__ void make_rots(double th, double ph, R* fwd, R* inv) {
    const double c = cos(0.5 * th), s = sin(0.5 * th);
    const double cp = cos(0.5 * ph), sp = sin(0.5 * ph);
    // RZ RY = [[e^{-i ph/2} c, -e^{-i ph/2} s], [e^{i ph/2} s, e^{i ph/2} c]]
    const double f[8] = {cp * c, -sp * c, -cp * s, sp * s, cp * s, sp * s, cp * c, sp * c};
    for (int i = 0; i < 8; ++i) fwd[i] = (R)f[i];
    // conjugate transpose: entry (r, c) = conj(f(c, r))
    const double g[8] = {f[0], -f[1], f[4], -f[5], f[2], -f[3], f[6], -f[7]};
    for (int i = 0; i < 8; ++i) inv[i] = (R)g[i];
}

// One block per circuit: rots[(param_off / 2 + n*l + q) * 24 + (0..7 | 8..15)]
// = (RZ RY, (RZ RY)^dagger) of qubit q in layer l (16..23 unused).
template <typename R>
__global__ void k_prepare_rots(const BatchCircuit* __restrict__ circs, const double* __restrict__ params,
                               R* __restrict__ rots) {
    const BatchCircuit C = circs[blockIdx.x];
    const int n = C.n;
    for (int k = threadIdx.x; k < n * C.layers; k += blockDim.x) {
        const int l = k / n, q = k % n;
        const double* P = params + C.param_off + 2 * n * l;
        R* o = rots + (C.param_off / 2 + k) * 24;
        make_rots<R>(P[q], P[n + q], o, o + 8);
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
// for complex64, 2..3 for complex128.
template <typename R, int GW>
__global__ void __launch_bounds__(512) k_batch_hea(const BatchCircuit* __restrict__ circs, int ncirc,
                                                   const R* __restrict__ rots, const double* __restrict__ params,
                                                   double* __restrict__ grads,
                                                   double* __restrict__ energies, cv<R>* __restrict__ scratch,
                                                   uint64_t scratch_elems, int* __restrict__ queue, double hx, int bmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = T >> 6;
    cv<R>* tp = reinterpret_cast<cv<R>*>(smem);
    cv<R>* tl = tp + (1u << bmax);
    double* wred = reinterpret_cast<double*>(tl + (1u << bmax));  // [nw][kMaxSlots]
    __shared__ int s_next;
    cv<R>* gpsi = scratch + (uint64_t)blockIdx.x * 2 * scratch_elems;
    cv<R>* glam = gpsi + scratch_elems;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_next = atomicAdd(queue, 1);
        __syncthreads();
        const int ci = s_next;
        if (ci >= ncirc) break;
        const BatchCircuit C = circs[ci];
        const int n = C.n, L = C.layers;
        const int b = min(C.tile_bits, n);
        const bool resident = (n <= b);
        const uint32_t TS = 1u << b;
        const uint64_t N = 1ULL << n;
        const R* RT = rots + (C.param_off / 2) * 24;  // [l][q][24]
        double* G = grads + C.param_off;
        const uint64_t ntA = resident ? 1 : (N >> b);
        const int h = n - b;

        // LDS position of tile element j: the low four index bits XOR the next
        // four, so sixteen lanes reading sixteen elements of a register group
        // hit distinct 8-byte slots whichever tile bits the group holds
        auto P = [](uint32_t j) __attribute__((always_inline)) -> uint32_t { return j ^ ((j >> 4) & 15u); };
        // qubits per register group, forward and reverse (sub-cubes of 2^GW
        // elements; the default is 4 for complex64 and 3 for complex128)
        constexpr int GM = GW;
        constexpr int GA = GW;
        // --- tile transfer helpers: eight elements per thread in flight before
        // any is written on, so a tile costs about one memory latency, not
        // one per element
        // psi (and lambda) tile in one round of loads: up to 16 elements per
        // thread per array in flight (64 KiB per CU for a 2^13 complex64 tile)
        auto load2 = [&](int kind, uint64_t t, bool two) __attribute__((always_inline)) {
            for (uint32_t j0 = tid; j0 < TS; j0 += 16 * T) {
                cv<R> v[16], w[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {  // every element assigned: the arrays stay in registers
                    const uint64_t gi = tile_global(kind, t, min(j0 + u * T, TS - 1), b, n);
                    v[u] = gpsi[gi];
                    w[u] = two ? glam[gi] : v[u];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const uint32_t j = j0 + u * T;
                    if (j < TS) {
                        tp[P(j)] = v[u];
                        if (two) tl[P(j)] = w[u];
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
        auto store = [&](int kind, uint64_t t, const cv<R>* src, cv<R>* dst) __attribute__((always_inline)) {
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
                    if (j < TS) dst[tile_global(kind, t, j, b, n)] = v[u];
                }
            }
        };
        // Rotations of tile bits [t0, t0 + G) (qubit = tile bit + qoff): one LDS
        // round trip per group, not per qubit.  Each thread takes sub-cubes of
        // 2^G elements that differ only in those bits into registers and
        // applies the G rotations there.  G is a compile-time constant so the
        // sub-cube stays in registers.
        auto rot_g = [&](auto Gc, cv<R>* tile, int t0, int qoff, int l) __attribute__((always_inline)) {
            constexpr int G = decltype(Gc)::value;
            constexpr uint32_t D = 1u << G;
            const uint32_t cubes = TS >> G, lo = (1u << t0) - 1;
            for (uint32_t s = tid; s < cubes; s += T) {
                const uint32_t base = ((s & ~lo) << G) | (s & lo);
                const uint32_t pb0 = P(base);  // P is linear over disjoint bits
                cv<R> x[D];
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) x[m] = tile[pb0 ^ P(m << t0)];
#pragma unroll
                for (int jb = 0; jb < G; ++jb) {
                    const R* u = RT + (n * l + (t0 + jb + qoff)) * 24;
#pragma unroll
                    for (uint32_t m = 0; m < D; ++m)
                        if (!(m & (1u << jb))) mul2<R>(u, x[m].x, x[m].y, x[m | (1u << jb)].x, x[m | (1u << jb)].y);
                }
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) tile[pb0 ^ P(m << t0)] = x[m];
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
        auto rot_group = [&](cv<R>* tile, int t0, int g, int qoff, int l) __attribute__((always_inline)) {
            static_assert(GM >= 2 && GM <= 4, "group width");
            switch (g) {
                case 1: rot_g(std::integral_constant<int, 1>{}, tile, t0, qoff, l); break;
                case 2: rot_g(std::integral_constant<int, 2>{}, tile, t0, qoff, l); break;
                case 3: if constexpr (GM >= 3) rot_g(std::integral_constant<int, 3>{}, tile, t0, qoff, l); break;
                case 4: if constexpr (GM >= 4) rot_g(std::integral_constant<int, 4>{}, tile, t0, qoff, l); break;
                default: break;
            }
        };
        // Reverse steps of tile bits [t0, t0 + G) on psi and lambda.  For qubit
        // q with U = RZ(phi) RY(theta), both gradients are read before undoing U:
        //   dE/dphi   = Im<lambda|Z psi>  = sum Im(l0* p0) - Im(l1* p1)
        //   dE/dtheta = Im<lambda|RZ Y RZ^dag psi>,  RZ Y RZ^dag = [[0, -i e^{-i phi}], [i e^{i phi}, 0]]
        //             = cos(phi) A - sin(phi) B,  A = sum Re(l1* p0) - Re(l0* p1),
        //                                         B = sum Im(l1* p0) + Im(l0* p1)
        // then one (RZ RY)^dagger on psi and lambda.  Products are summed in the
        // state's precision over a sub-cube, then in FP64; cos and sin are applied
        // in FP64 when the layer's gradients are written.  Rotations on different
        // qubits commute, so the qubits of a group may go in any order.
        auto adj_g = [&](auto Gc, int t0, int qoff, int l) __attribute__((always_inline)) {
            constexpr int G = decltype(Gc)::value;
            constexpr uint32_t D = 1u << G;
            const uint32_t cubes = TS >> G, lo = (1u << t0) - 1;
            double az[G], aa[G], ab[G];
#pragma unroll
            for (int jb = 0; jb < G; ++jb) az[jb] = aa[jb] = ab[jb] = 0.0;
            for (uint32_t s = tid; s < cubes; s += T) {
                const uint32_t base = ((s & ~lo) << G) | (s & lo);
                const uint32_t pb0 = P(base);  // P is linear over disjoint bits
                cv<R> x[D], y[D];
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) {
                    x[m] = tp[pb0 ^ P(m << t0)];
                    y[m] = tl[pb0 ^ P(m << t0)];
                }
                // every qubit's products from the same registers (rotations on
                // the group's other qubits, applied to both, leave them
                // unchanged); Im<lambda|Z psi> from one product per element:
                // sum over bit = 0 minus bit = 1 of w_m = Im(conj(l_m) p_m)
                R w[D], wall = 0;
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) {
                    w[m] = y[m].x * x[m].y - y[m].y * x[m].x;
                    wall += w[m];
                }
#pragma unroll
                for (int jb = G - 1; jb >= 0; --jb) {
                    const uint32_t bit = 1u << jb;
                    R w1 = 0, pa = 0, pb = 0;
#pragma unroll
                    for (uint32_t m = 0; m < D; ++m) {
                        if (m & bit) continue;
                        const cv<R>&a0 = x[m], &a1 = x[m | bit], &b0 = y[m], &b1 = y[m | bit];
                        w1 += w[m | bit];
                        pa += (b1.x * a0.x + b1.y * a0.y) - (b0.x * a1.x + b0.y * a1.y);
                        pb += (b1.x * a0.y - b1.y * a0.x) + (b0.x * a1.y - b0.y * a1.x);
                    }
                    az[jb] += (double)(wall - 2 * w1);
                    aa[jb] += (double)pa;
                    ab[jb] += (double)pb;
                }
#pragma unroll
                for (int jb = G - 1; jb >= 0; --jb) {
                    const R* inv = RT + (n * l + (t0 + jb + qoff)) * 24 + 8;
                    const uint32_t bit = 1u << jb;
#pragma unroll
                    for (uint32_t m = 0; m < D; ++m) {
                        if (m & bit) continue;
                        mul2<R>(inv, x[m].x, x[m].y, x[m | bit].x, x[m | bit].y);
                        mul2<R>(inv, y[m].x, y[m].y, y[m | bit].x, y[m | bit].y);
                    }
                }
#pragma unroll
                for (uint32_t m = 0; m < D; ++m) {
                    tp[pb0 ^ P(m << t0)] = x[m];
                    tl[pb0 ^ P(m << t0)] = y[m];
                }
            }
            // wave sums of the 3G products (qubit jb: z, a, b at 3 jb + 0..2),
            // reduce-scatter: offsets 32 and 16 halve the values a lane
            // carries (12 -> 6 -> 3), offsets 8..1 finish each of the 3
            {
                double acc[12];
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    const int jb = k / 3;
                    acc[k] = jb >= G ? 0.0 : (k % 3 == 0 ? az[jb < G ? jb : 0] : k % 3 == 1 ? aa[jb < G ? jb : 0] : ab[jb < G ? jb : 0]);
                }
                const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0;
                double v6[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) v6[i] = (h5 ? acc[6 + i] : acc[i]) + __shfl_xor(h5 ? acc[i] : acc[6 + i], 32, 64);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    double xs = (h4 ? v6[3 + i] : v6[i]) + __shfl_xor(h4 ? v6[i] : v6[3 + i], 16, 64);
#pragma unroll
                    for (int off = 8; off > 0; off >>= 1) xs += __shfl_xor(xs, off, 64);
                    const int k = (h5 ? 6 : 0) + (h4 ? 3 : 0) + i;
                    if ((lane & 15) == 0 && k < 3 * G) wred[wave * kMaxSlots + 3 * (t0 + qoff) + k] += xs;
                }
            }
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
        // and back.  (Prefetching the next tile into registers across `body`
        // needs 64 more VGPRs per lane and spilled at 512 threads.)
        auto tile_pass = [&](int kind, bool two, auto&& body) __attribute__((always_inline)) {
            for (uint64_t t = 0; t < ntA; ++t) {
                load2(kind, t, two);
                __syncthreads();
                body(t);
                store(kind, t, tp, gpsi);
                if (two) store(kind, t, tl, glam);
                __syncthreads();
            }
        };
        auto cz = [&](int kind, uint64_t t, bool both) __attribute__((always_inline)) {
            for (uint32_t j = tid; j < TS; j += T) {
                if (cz_ring_sign(tile_global(kind, t, j, b, n), n)) {
                    cv<R> v = tp[P(j)];
                    v.x = -v.x; v.y = -v.y;
                    tp[P(j)] = v;
                    if (both) {
                        cv<R> w = tl[P(j)];
                        w.x = -w.x; w.y = -w.y;
                        tl[P(j)] = w;
                    }
                }
            }
            __syncthreads();
        };
        auto clear_slots = [&]() __attribute__((always_inline)) {
            for (int i = tid; i < nw * kMaxSlots; i += T) wred[i] = 0.0;
            __syncthreads();
        };
        // layer gradients from the slots (3q: dE/dphi_q, 3q+1 / 3q+2: A, B)
        auto flush_slots = [&](int l) __attribute__((always_inline)) {
            for (int q = tid; q < n; q += T) {
                double z = 0.0, va = 0.0, vb = 0.0;
                for (int w = 0; w < nw; ++w) {
                    z += wred[w * kMaxSlots + 3 * q];
                    va += wred[w * kMaxSlots + 3 * q + 1];
                    vb += wred[w * kMaxSlots + 3 * q + 2];
                }
                const double ph = params[C.param_off + 2 * n * l + n + q];
                G[2 * n * l + n + q] = z;
                G[2 * n * l + q] = cos(ph) * va - sin(ph) * vb;
            }
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
        for (int l = 0; l < L; ++l) {
            if (resident) {
                for_groups(0, n, GM, [&](int t0, int g) { rot_group(tp, t0, g, 0, l); });
                cz(0, 0, false);
            } else {
                tile_pass(1, false, [&](uint64_t) {  // pass A: qubits 0..b-1
                    for_groups(0, b, GM, [&](int t0, int g) { rot_group(tp, t0, g, 0, l); });
                });
                tile_pass(2, false, [&](uint64_t t) {  // pass B: qubits b..n-1, then the CZ ring
                    for_groups(b - h, b, GM, [&](int t0, int g) { rot_group(tp, t0, g, h, l); });
                    cz(2, t, false);
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
                store(2, t, tl, glam);
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
        for (int l = L - 1; l >= 0; --l) {
            clear_slots();
            if (resident) {
                cz(0, 0, true);
                for_groups(0, n, GA, [&](int t0, int g) { adj_group(t0, g, 0, l); });
            } else {
                tile_pass(2, true, [&](uint64_t t) {  // pass B': CZ ring, qubits n-1..b
                    cz(2, t, true);
                    for_groups(b - h, b, GA, [&](int t0, int g) { adj_group(t0, g, h, l); });
                });
                tile_pass(1, true, [&](uint64_t) {  // pass A': qubits b-1..0
                    for_groups(0, b, GA, [&](int t0, int g) { adj_group(t0, g, 0, l); });
                });
            }
            flush_slots(l);
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
    int bused = 6;  // widest tile of the batch: the LDS each workgroup needs
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
        bused = std::max(bused, std::min(tb, nq[c] > tb ? tb : std::max(nq[c], 6)));
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
    const size_t lds = 2 * ((size_t)1 << bused) * elem + (size_t)(T / 64) * kMaxSlots * sizeof(double);
    const int gw = B.group_width ? B.group_width : iq_batch_max_group_width(B.prec);
    const void* kern = B.prec == kC128 ? (gw == 2 ? (const void*)k_batch_hea<double, 2> : (const void*)k_batch_hea<double, 3>)
                       : gw == 2       ? (const void*)k_batch_hea<float, 2>
                       : gw == 3       ? (const void*)k_batch_hea<float, 3>
                                       : (const void*)k_batch_hea<float, 4>;
    hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, T, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    const int nblk = num_blocks > 0 ? num_blocks : std::max(1, cus * per_cu);
    const uint64_t scratch_elems = nmax ? (1ULL << nmax) : 1;
    if (!grow(reinterpret_cast<void**>(&B.circs), B.circ_cap, ncirc * sizeof(BatchCircuit)) ||
        !grow(reinterpret_cast<void**>(&B.params), B.param_cap, (size_t)total * 8) ||
        !grow(reinterpret_cast<void**>(&B.energies), B.energy_cap, ncirc * 8) ||
        !grow(&B.scratch, B.scratch_bytes, (size_t)nblk * 2 * scratch_elems * elem) ||
        !grow(&B.rots, B.rots_cap, (size_t)(total / 2) * 24 * (B.prec == kC128 ? 8 : 4)))
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
                       B.params, B.grads, B.energies, (CV*)B.scratch, scratch_elems, B.queue, hx, bused)
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
