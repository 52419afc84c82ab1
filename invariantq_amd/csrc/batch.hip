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

constexpr int kMaxSlots = 64;  // gradient slots per layer (2 * n <= 40)

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

// RZ(phi) RY(theta) and the two daggers, from FP64 angles (24 values).
template <typename R>
__device__ __forceinline__ void make_rots(double th, double ph, R* fwd, R* rzd, R* ryd) {
    const double c = cos(0.5 * th), s = sin(0.5 * th);
    const double cp = cos(0.5 * ph), sp = sin(0.5 * ph);
    // RZ RY = [[e^{-i ph/2} c, -e^{-i ph/2} s], [e^{i ph/2} s, e^{i ph/2} c]]
    const double f[8] = {cp * c, -sp * c, -cp * s, sp * s, cp * s, sp * s, cp * c, sp * c};
    for (int i = 0; i < 8; ++i) fwd[i] = (R)f[i];
    // RZ^dag = diag(e^{i ph/2}, e^{-i ph/2})
    const double z[8] = {cp, sp, 0, 0, 0, 0, cp, -sp};
    for (int i = 0; i < 8; ++i) rzd[i] = (R)z[i];
    // RY^dag = [[c, s], [-s, c]]
    const double y[8] = {c, 0, s, 0, -s, 0, c, 0};
    for (int i = 0; i < 8; ++i) ryd[i] = (R)y[i];
}

// One block per circuit: rots[(param_off / 2 + n*l + q) * 24 + (0..7 | 8..15 | 16..23)]
// = (RZ RY, RZ^dag, RY^dag) of qubit q in layer l.
template <typename R>
__global__ void k_prepare_rots(const BatchCircuit* __restrict__ circs, const double* __restrict__ params,
                               R* __restrict__ rots) {
    const BatchCircuit C = circs[blockIdx.x];
    const int n = C.n;
    for (int k = threadIdx.x; k < n * C.layers; k += blockDim.x) {
        const int l = k / n, q = k % n;
        const double* P = params + C.param_off + 2 * n * l;
        R* o = rots + (C.param_off / 2 + k) * 24;
        make_rots<R>(P[q], P[n + q], o, o + 8, o + 16);
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

template <typename R>
__global__ void __launch_bounds__(512) k_batch_hea(const BatchCircuit* __restrict__ circs, int ncirc,
                                                   const R* __restrict__ rots, double* __restrict__ grads,
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

        // --- tile transfer helpers
        auto load = [&](int kind, uint64_t t, cv<R>* dst, const cv<R>* src) {
            for (uint32_t j = tid; j < TS; j += T) dst[j] = src[tile_global(kind, t, j, b, n)];
        };
        auto store = [&](int kind, uint64_t t, const cv<R>* src, cv<R>* dst) {
            for (uint32_t j = tid; j < TS; j += T) dst[tile_global(kind, t, j, b, n)] = src[j];
        };
        // rotation of tile bit tb by a 2x2 matrix on one tile
        auto rot = [&](cv<R>* tile, int tb, const R* u) {
            const uint32_t bit = 1u << tb, lo = bit - 1;
            for (uint32_t p = tid; p < (TS >> 1); p += T) {
                const uint32_t i0 = ((p & ~lo) << 1) | (p & lo), i1 = i0 | bit;
                cv<R> a = tile[i0], c = tile[i1];
                mul2<R>(u, a.x, a.y, c.x, c.y);
                tile[i0] = a;
                tile[i1] = c;
            }
            __syncthreads();
        };
        // reverse step on tile bit tb: grad phi (Z) then un-RZ, grad theta (Y) then un-RY
        auto adj = [&](int tb, const R* rzd, const R* ryd, int slot) {
            const uint32_t bit = 1u << tb, lo = bit - 1;
            double gz = 0.0, gy = 0.0;
            for (uint32_t p = tid; p < (TS >> 1); p += T) {
                const uint32_t i0 = ((p & ~lo) << 1) | (p & lo), i1 = i0 | bit;
                cv<R> a0 = tp[i0], a1 = tp[i1], b0 = tl[i0], b1 = tl[i1];
                gz += ((double)b0.x * a0.y - (double)b0.y * a0.x) - ((double)b1.x * a1.y - (double)b1.y * a1.x);
                mul2<R>(rzd, a0.x, a0.y, a1.x, a1.y);
                mul2<R>(rzd, b0.x, b0.y, b1.x, b1.y);
                gy += (-(double)b0.x * a1.x - (double)b0.y * a1.y) + ((double)b1.x * a0.x + (double)b1.y * a0.y);
                mul2<R>(ryd, a0.x, a0.y, a1.x, a1.y);
                mul2<R>(ryd, b0.x, b0.y, b1.x, b1.y);
                tp[i0] = a0; tp[i1] = a1; tl[i0] = b0; tl[i1] = b1;
            }
            gz = wave_sum(gz);
            gy = wave_sum(gy);
            if (lane == 0) {
                wred[wave * kMaxSlots + 2 * slot + 1] += gz;  // RZ parameter slot n + q
                wred[wave * kMaxSlots + 2 * slot] += gy;      // RY parameter slot q
            }
            __syncthreads();
        };
        auto cz = [&](int kind, uint64_t t, bool both) {
            for (uint32_t j = tid; j < TS; j += T) {
                if (cz_ring_sign(tile_global(kind, t, j, b, n), n)) {
                    cv<R> v = tp[j];
                    v.x = -v.x; v.y = -v.y;
                    tp[j] = v;
                    if (both) {
                        cv<R> w = tl[j];
                        w.x = -w.x; w.y = -w.y;
                        tl[j] = w;
                    }
                }
            }
            __syncthreads();
        };
        auto clear_slots = [&]() {
            for (int i = tid; i < nw * kMaxSlots; i += T) wred[i] = 0.0;
            __syncthreads();
        };
        // layer gradients: slot 2q = theta_q, 2q+1 = phi_q
        auto flush_slots = [&](int l) {
            for (int s = tid; s < 2 * n; s += T) {
                double v = 0.0;
                for (int w = 0; w < nw; ++w) v += wred[w * kMaxSlots + s];
                const int q = s >> 1;
                G[2 * n * l + ((s & 1) ? n + q : q)] = v;
            }
            __syncthreads();
        };

        // --- forward: psi = U|0>
        if (resident) {
            for (uint32_t j = tid; j < TS; j += T) {
                cv<R> v;
                v.x = (j == 0) ? (R)1 : (R)0;
                v.y = 0;
                tp[j] = v;
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
                for (int q = 0; q < n; ++q) rot(tp, q, RT + (n * l + q) * 24);
                cz(0, 0, false);
            } else {
                for (uint64_t t = 0; t < ntA; ++t) {  // pass A: qubits 0..b-1
                    load(1, t, tp, gpsi);
                    __syncthreads();
                    for (int q = 0; q < b; ++q) rot(tp, q, RT + (n * l + q) * 24);
                    store(1, t, tp, gpsi);
                    __syncthreads();
                }
                for (uint64_t t = 0; t < ntA; ++t) {  // pass B: qubits b..n-1, then the CZ ring
                    load(2, t, tp, gpsi);
                    __syncthreads();
                    for (int q = b; q < n; ++q) rot(tp, q - b + (b - h), RT + (n * l + q) * 24);
                    cz(2, t, false);
                    store(2, t, tp, gpsi);
                    __syncthreads();
                }
            }
        }

        // --- lambda = H psi, energy = Re <psi|lambda>
        double e = 0.0;
        auto hdiag = [&](uint64_t g) {
            const uint64_t diff = (g ^ (g >> 1)) & ((1ULL << (n - 1)) - 1);
            return (double)(n - 1) - 2.0 * (double)__popcll(diff);
        };
        if (resident) {
            for (uint32_t j = tid; j < TS; j += T) {
                const cv<R> v = tp[j];
                double sr = hdiag(j) * (double)v.x, si = hdiag(j) * (double)v.y;
                for (int q = 0; q < n; ++q) {
                    const cv<R> w = tp[j ^ (1u << q)];
                    sr += hx * (double)w.x;
                    si += hx * (double)w.y;
                }
                cv<R> o;
                o.x = (R)sr;
                o.y = (R)si;
                tl[j] = o;
                e += (double)v.x * sr + (double)v.y * si;
            }
            __syncthreads();
        } else {
            for (uint64_t t = 0; t < ntA; ++t) {  // diagonal + X on the b low qubits
                load(1, t, tp, gpsi);
                __syncthreads();
                for (uint32_t j = tid; j < TS; j += T) {
                    const uint64_t g = tile_global(1, t, j, b, n);
                    const cv<R> v = tp[j];
                    double sr = hdiag(g) * (double)v.x, si = hdiag(g) * (double)v.y;
                    for (int q = 0; q < b; ++q) {
                        const cv<R> w = tp[j ^ (1u << q)];
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
                    const cv<R> v = tp[j];
                    cv<R> o = tl[j];
                    double sr = (double)o.x, si = (double)o.y;
                    for (int q = b; q < n; ++q) {
                        const cv<R> w = tp[j ^ (1u << (q - b + (b - h)))];
                        sr += hx * (double)w.x;
                        si += hx * (double)w.y;
                    }
                    o.x = (R)sr;
                    o.y = (R)si;
                    tl[j] = o;
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
                for (int q = n - 1; q >= 0; --q) adj(q, RT + (n * l + q) * 24 + 8, RT + (n * l + q) * 24 + 16, q);
            } else {
                for (uint64_t t = 0; t < ntA; ++t) {  // pass B': CZ ring, qubits n-1..b
                    load(2, t, tp, gpsi);
                    load(2, t, tl, glam);
                    __syncthreads();
                    cz(2, t, true);
                    for (int q = n - 1; q >= b; --q)
                        adj(q - b + (b - h), RT + (n * l + q) * 24 + 8, RT + (n * l + q) * 24 + 16, q);
                    store(2, t, tp, gpsi);
                    store(2, t, tl, glam);
                    __syncthreads();
                }
                for (uint64_t t = 0; t < ntA; ++t) {  // pass A': qubits b-1..0
                    load(1, t, tp, gpsi);
                    load(1, t, tl, glam);
                    __syncthreads();
                    for (int q = b - 1; q >= 0; --q) adj(q, RT + (n * l + q) * 24 + 8, RT + (n * l + q) * 24 + 16, q);
                    store(1, t, tp, gpsi);
                    store(1, t, tl, glam);
                    __syncthreads();
                }
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
    for (size_t c = 0; c < ncirc; ++c) {
        if (nq[c] < 2 || nq[c] > 30 || layers[c] < 1 || 2 * nq[c] > kMaxSlots) {
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
    const int nblk = num_blocks > 0 ? num_blocks : std::max(1, cus);
    const size_t elem = B.prec == kC128 ? 16 : 8;
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
    const size_t lds = 2 * ((size_t)1 << bmax) * elem + (size_t)(T / 64) * kMaxSlots * sizeof(double);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, st);
    if (B.prec == kC128) {
        hipFuncSetAttribute((const void*)k_batch_hea<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_batch_hea<double>, dim3(nblk), dim3(T), lds, st, B.circs, (int)ncirc, (const double*)B.rots, B.grads,
                           B.energies, (double2*)B.scratch, scratch_elems, B.queue, hx, bmax);
    } else {
        hipFuncSetAttribute((const void*)k_batch_hea<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_batch_hea<float>, dim3(nblk), dim3(T), lds, st, B.circs, (int)ncirc, (const float*)B.rots, B.grads,
                           B.energies, (float2*)B.scratch, scratch_elems, B.queue, hx, bmax);
    }
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
