// Element-wise passes, FP64 reductions, sampling and partial trace.
//
// Reductions accumulate in FP64 whatever the amplitude precision, write one
// partial per block, and a single-block finaliser sums the partials in a fixed
// order: results are bit-reproducible run to run (no atomics on the result path).

#include "jacobi_rows.hpp"
#include "kernels.hpp"
#include "memio.hpp"
#include "reduce.hpp"

#include <algorithm>

namespace iqg {

namespace {

// blocks for `items` threads, at most `cap` and below 2^32 threads in all
// (every kernel launched with it loops over its items)
inline unsigned blocks_for(uint64_t items, unsigned block, uint64_t cap) {
    uint64_t g = (items + block - 1) / block;
    g = std::max<uint64_t>(1, std::min<uint64_t>({g, cap, kMaxDispatchItems / block}));
    return (unsigned)g;
}

// Whole-state passes: a grid-stride loop with four independent elements per
// thread and iteration (four loads in flight), then a tail of single ones.
// NT adds the streaming hint (states larger than kCacheStateMax).
template <int U, typename Body>
__device__ __forceinline__ void grid_stride(uint64_t count, Body&& body) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < count; i += U * stride) body(i, stride, U);
    for (; i < count; i += stride) body(i, stride, 1);
}

// Read+write passes: a block takes spans of U * blockDim consecutive
// elements (element span + u * blockDim + tid), so a wave's U accesses fall
// in one span instead of U regions a grid apart; the body sees the same
// (first element, stride, count) triple as with grid_stride.
template <int U, typename Body>
__device__ __forceinline__ void block_spans(uint64_t count, Body&& body) {
    const uint64_t span = (uint64_t)U * blockDim.x;
    for (uint64_t b = (uint64_t)blockIdx.x * span; b < count; b += (uint64_t)gridDim.x * span) {
        const uint64_t i = b + threadIdx.x;
        if (b + span <= count) {
            body(i, (uint64_t)blockDim.x, U);
        } else if (i < count) {
            body(i, (uint64_t)blockDim.x, (int)std::min<uint64_t>(U, (count - i + blockDim.x - 1) / blockDim.x));
        }
    }
}

template <typename R, bool NT>
__global__ void __launch_bounds__(256) k_init_basis(cv<R>* __restrict__ psi, uint64_t count, uint64_t index) {
    block_spans<kPassSpan>(count, [&](uint64_t i, uint64_t stride, int n) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u >= n) break;
            const uint64_t j = i + u * stride;
            cv<R> v;
            v.x = (j == index) ? (R)1 : (R)0;
            v.y = 0;
            st_stream<R>(psi + j, v, NT);
        }
    });
}

// Amplitudes whose target bit is the outcome are scaled; the others are
// written as zeros without being read, so the pass moves the kept half twice
// and the discarded half once.
template <typename R, bool NT>
__global__ void __launch_bounds__(256) k_collapse(cv<R>* __restrict__ psi, uint64_t count, int target, int outcome,
                                                  double factor) {
    const R f = (R)factor;
    block_spans<kPassSpan>(count, [&](uint64_t i, uint64_t stride, int n) {
        cv<R> v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < n && (int)(((i + u * stride) >> target) & 1ULL) == outcome) v[u] = ld_stream<R>(psi + i + u * stride, NT);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u >= n) break;
            const uint64_t j = i + u * stride;
            if ((int)((j >> target) & 1ULL) == outcome) {
                v[u].x *= f;
                v[u].y *= f;
            } else {
                v[u].x = 0;
                v[u].y = 0;
            }
            st_stream<R>(psi + j, v[u], NT);
        }
    });
}

template <typename R, bool NT>
__global__ void __launch_bounds__(256) k_scale(cv<R>* __restrict__ psi, uint64_t count, double factor) {
    const R f = (R)factor;
    block_spans<kPassSpan>(count, [&](uint64_t i, uint64_t stride, int n) {
        cv<R> v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < n) v[u] = ld_stream<R>(psi + i + u * stride, NT);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u >= n) break;
            v[u].x *= f;
            v[u].y *= f;
            st_stream<R>(psi + i + u * stride, v[u], NT);
        }
    });
}

// FP64 copies of amplitudes: re[i * es], im[i * es] (es = 1: split arrays;
// es = 2 with im = re + 1: interleaved complex)
template <typename R>
__global__ void __launch_bounds__(256) k_split(const cv<R>* __restrict__ psi, uint64_t count, double* re, double* im,
                                               uint32_t es) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const cv<R> v = psi[i];
        re[i * es] = (double)v.x;
        im[i * es] = (double)v.y;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_join(cv<R>* __restrict__ psi, uint64_t count, const double* re,
                                              const double* im, uint32_t es) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        cv<R> v;
        v.x = (R)re[i * es];
        v.y = (R)im[i * es];
        psi[i] = v;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_probs(const cv<R>* __restrict__ psi, uint64_t count, double* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const cv<R> v = psi[i];
        const double x = (double)v.x, y = (double)v.y;
        out[i] = x * x + y * y;
    }
}

// ------------------------------------------------------------ reductions

__global__ void __launch_bounds__(256) k_finalize(const double* __restrict__ part, int nblocks, int m_used,
                                                  double* __restrict__ out) {
    __shared__ double sh[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int m = 0; m < m_used; ++m) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += part[(uint64_t)b * m_used + m];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (lane == 0) sh[wave] = s;
        __syncthreads();
        if (threadIdx.x == 0) out[m] = sh[0] + sh[1] + sh[2] + sh[3];
        __syncthreads();
    }
}

template <typename R, bool NT>
__global__ void __launch_bounds__(256) k_red_norm(const cv<R>* __restrict__ psi, uint64_t count, double* __restrict__ part) {
    double v[1] = {0.0};
    grid_stride<4>(count, [&](uint64_t i, uint64_t stride, int n) {
        cv<R> a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < n) a[u] = ld_stream<R>(psi + i + u * stride, NT);
#pragma unroll
        for (int u = 0; u < 4; ++u) {  // in element order, as a single-element loop would add them
            if (u >= n) break;
            const double x = (double)a[u].x, y = (double)a[u].y;
            v[0] += x * x + y * y;
        }
    });
    block_sum_store<1>(v, 1, part);
}

template <typename R, bool NT>
__global__ void __launch_bounds__(256) k_red_bitprob(const cv<R>* __restrict__ psi, uint64_t count, int target,
                                                     double* __restrict__ part) {
    double v[2] = {0.0, 0.0};
    grid_stride<4>(count, [&](uint64_t i, uint64_t stride, int n) {
        cv<R> a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u < n) a[u] = ld_stream<R>(psi + i + u * stride, NT);
#pragma unroll
        for (int u = 0; u < 4; ++u) {  // in element order, as a single-element loop would add them
            if (u >= n) break;
            const double x = (double)a[u].x, y = (double)a[u].y;
            const double p = x * x + y * y;
            if (((i + u * stride) >> target) & 1ULL) v[1] += p;
            else v[0] += p;
        }
    });
    block_sum_store<2>(v, 2, part);
}

constexpr int kPauliPerGroup = 8;

struct PauliGroup {
    uint64_t xmask;
    uint64_t zy[kPauliPerGroup];
    int ny[kPauliPerGroup];
    int nstr;
};

// <psi|P|psi> = sum_i conj(psi[i ^ x]) * i^{nY} * (-1)^{popcount(i & (Y|Z))} * psi[i],
// skipping |psi_i|^2 < 1e-18 (reference core/src/statevector.cpp:265-288).
// With a flip mask x, index j (bit hbit, the highest of x, clear) and its
// partner j ^ x are taken together, so every amplitude is read once; each of
// the two still adds its own term, under its own skip test.  Four pairs (or
// four amplitudes, x = 0) per thread and iteration are loaded at once.
template <typename R, bool NT>
__global__ void __launch_bounds__(256) k_red_pauli(const cv<R>* __restrict__ psi, uint64_t count, PauliGroup g,
                                                   int hbit, double* __restrict__ part) {
    double v[2 * kPauliPerGroup];
#pragma unroll
    for (int m = 0; m < 2 * kPauliPerGroup; ++m) v[m] = 0.0;
    // term of index i: conj(b) * phase_i * a with a = psi[i], b = psi[i ^ x]
    auto term = [&](uint64_t i, const cv<R>& a, const cv<R>& b) __attribute__((always_inline)) {
        const double ar = (double)a.x, ai = (double)a.y;
        if (ar * ar + ai * ai < 1e-18) return;
        const double br = (double)b.x, bi = (double)b.y;
        const double pr = br * ar + bi * ai;
        const double pi = br * ai - bi * ar;
#pragma unroll
        for (int s = 0; s < kPauliPerGroup; ++s) {
            if (s >= g.nstr) break;
            const double sg = (__popcll(i & g.zy[s]) & 1) ? -1.0 : 1.0;
            double cr = sg * pr, ci = sg * pi;
            switch (g.ny[s] & 3) {
                case 1: { const double t = cr; cr = -ci; ci = t; break; }    // * i
                case 2: cr = -cr; ci = -ci; break;                           // * -1
                case 3: { const double t = cr; cr = ci; ci = -t; break; }    // * -i
                default: break;
            }
            v[2 * s] += cr;
            v[2 * s + 1] += ci;
        }
    };
    if (g.xmask == 0) {
        grid_stride<4>(count, [&](uint64_t i, uint64_t stride, int n) {
            cv<R> a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u < n) a[u] = ld_stream<R>(psi + i + u * stride, NT);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u < n) term(i + u * stride, a[u], a[u]);
        });
    } else {
        const uint64_t lo = (1ULL << hbit) - 1;
        grid_stride<4>(count >> 1, [&](uint64_t k, uint64_t stride, int n) {
            uint64_t j[4];
            cv<R> a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u < n) {
                    const uint64_t kk = k + u * stride;
                    j[u] = ((kk & ~lo) << 1) | (kk & lo);
                    a[u] = ld_stream<R>(psi + j[u], NT);
                    b[u] = ld_stream<R>(psi + (j[u] ^ g.xmask), NT);
                }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u < n) {
                    term(j[u], a[u], b[u]);
                    term(j[u] ^ g.xmask, b[u], a[u]);
                }
        });
    }
    block_sum_store<2 * kPauliPerGroup>(v, 2 * g.nstr, part);
}

__global__ void __launch_bounds__(256) k_red_dm_diag(const double2* __restrict__ rho, int nq, int mode, int target,
                                                     double* __restrict__ part) {
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = 0.0;
    const uint64_t dim = 1ULL << nq;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const double d = rho[i * (dim + 1)].x;
        if (mode == 0) {
            v[0] += d;
        } else if (mode == 1) {
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (q < nq) v[q] += ((i >> q) & 1ULL) ? -d : d;
        } else if (mode == 2) {
            if ((i >> target) & 1ULL) v[1] += d;
            else v[0] += d;
        } else {
            const double p = d > 0.0 ? d : 0.0;
            if (p > 1e-14) v[0] -= p * log2(p);
        }
    }
    const int m_used = mode == 1 ? nq : (mode == 2 ? 2 : 1);
    block_sum_store<16>(v, m_used, part);
}

__global__ void __launch_bounds__(256) k_dm_collapse(double2* __restrict__ rho, int nq, int target, int outcome,
                                                     double factor) {
    const uint64_t count = 1ULL << (2 * nq);
    const uint64_t cmask = (1ULL << nq) - 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint64_t r = i >> nq, c = i & cmask;
        double2 v = rho[i];
        if ((int)((r >> target) & 1ULL) == outcome && (int)((c >> target) & 1ULL) == outcome) {
            v.x *= factor;
            v.y *= factor;
        } else {
            v.x = 0.0;
            v.y = 0.0;
        }
        rho[i] = v;
    }
}

// --------------------------------------------------------------- scan

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint64_t kScanChunk = (uint64_t)kScanThreads * kScanItems;

__device__ __forceinline__ double block_exclusive_scan(double x, double* total) {
    __shared__ double sh[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    double before = 0.0;
    for (int w = 0; w < wave; ++w) before += sh[w];
    if (total) *total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return before + incl - x;
}

// The CDF straight from the amplitudes, with no probability buffer: chunk
// sums of |psi_i|^2 (read only), an exclusive scan of the chunk sums, then
// either each chunk's prefix sums plus its offset (read the state again,
// write the CDF) or, for few shots, the prefix sums of only the chunks the
// shots land in.  Every pass stages a chunk's probabilities in LDS from
// coalesced loads; thread t then adds its kScanItems consecutive entries in
// index order and the block scans the thread sums, the same arithmetic in
// every pass (chunk_prefixes).  Fixed order throughout: the CDF is the same
// bits on every run.
//
// The parallel sums round differently from a serial running sum, so the raw
// values can step back by an ulp where one thread's or chunk's offset meets
// the next.  The CDF written is the running maximum of the raw values
// (monotone, and still within the rounding bound of the serial CDF), and a
// shot's index is the first whose raw value reaches the uniform, which the
// running maximum does not change.  The written-CDF search and the per-chunk
// search therefore return the same index for every uniform.
__device__ __forceinline__ uint32_t scan_lds(uint32_t i) { return i + (i >> 4); }  // one pad per 16 (banks)

template <typename R>
__device__ __forceinline__ void chunk_probs(const cv<R>* __restrict__ psi, uint64_t count, uint64_t base, double* sh) {
    cv<R> v[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kScanThreads + threadIdx.x;
        v[k] = psi[i < count ? i : count - 1];
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t j = (uint32_t)k * kScanThreads + threadIdx.x;
        const double x = (double)v[k].x, y = (double)v[k].y;
        sh[scan_lds(j)] = (base + j < count) ? fma(x, x, y * y) : 0.0;
    }
}

// Thread t's running sums over chunk entries t*kScanItems .. +kScanItems-1
// (loc) and its exclusive offset within the chunk (returned); `total` gets
// the chunk sum.  Ends with a barrier: sh may be rewritten.
template <typename R>
__device__ __forceinline__ double chunk_prefixes(const cv<R>* __restrict__ psi, uint64_t count, uint64_t base, double* sh,
                                                 double (&loc)[kScanItems], double* total) {
    chunk_probs<R>(psi, count, base, sh);
    __syncthreads();
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        run += sh[scan_lds(threadIdx.x * kScanItems + k)];
        loc[k] = run;
    }
    return block_exclusive_scan(run, total);
}

// entries of this thread's segment that lie below count
__device__ __forceinline__ int chunk_valid(uint64_t count, uint64_t base) {
    const uint64_t first = base + (uint64_t)threadIdx.x * kScanItems;
    return first >= count ? 0 : (int)std::min<uint64_t>(count - first, kScanItems);
}

// Exclusive running maximum of non-negative x over the block's threads (0
// for thread 0); `all` gets the block maximum.  Ends with a barrier.
__device__ __forceinline__ double block_exclusive_max(double x, double* all) {
    __shared__ double sh[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(incl, off, 64);
        if (lane >= off) incl = fmax(incl, y);
    }
    const double prev = __shfl_up(incl, 1, 64);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    double before = lane > 0 ? prev : 0.0;
    for (int w = 0; w < wave; ++w) before = fmax(before, sh[w]);
    if (all) *all = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    __syncthreads();
    return before;
}

// Pass 1: chunk sums (totals) and each chunk's largest in-chunk prefix at an
// index below count (tops; within a thread the prefixes do not decrease, so
// that is a thread's last one).
template <typename R>
__global__ void __launch_bounds__(kScanThreads) k_cdf_sums(const cv<R>* __restrict__ psi, uint64_t count,
                                                           double* __restrict__ totals, double* __restrict__ tops) {
    __shared__ double sh[kScanChunk + kScanChunk / 16];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    double loc[kScanItems];
    double total = 0.0;
    const double off = chunk_prefixes<R>(psi, count, base, sh, loc, &total);
    const int nv = chunk_valid(count, base);
    double top = 0.0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (k < nv) top = off + loc[k];
    double chunk_top = 0.0;
    block_exclusive_max(top, &chunk_top);
    if (threadIdx.x == 0) {
        totals[blockIdx.x] = total;
        tops[blockIdx.x] = chunk_top;
    }
}

// Pass 2 turns the chunk sums into exclusive offsets, in place, and tops[c]
// into the running maximum, over chunks 0..c, of each chunk's largest CDF
// value (top + offset).  Chunks are grouped 4096 to a block (thread t takes
// 16 consecutive chunks):
//   k_totals_sums   each group's sum (bsum)
//   k_scan_small    exclusive scan of the group sums (boff)
//   k_totals_apply  offsets and the running maximum within each group, and
//                   each group's maximum (bmax)
//   k_scan_small    exclusive running maximum of the group maxima (bpre)
// The running maximum at chunk c is then max(bpre[c / 4096], tops[c])
// (chunk_top).
constexpr uint64_t kGroupChunks = kScanChunk;

__device__ __forceinline__ double chunk_top(const double* __restrict__ tops, const double* __restrict__ bpre, uint64_t c) {
    return fmax(bpre[c / kGroupChunks], tops[c]);
}

__global__ void __launch_bounds__(kScanThreads) k_totals_sums(const double* __restrict__ totals, uint64_t nchunks,
                                                              double* __restrict__ bsum) {
    const uint64_t first = (uint64_t)blockIdx.x * kGroupChunks + (uint64_t)threadIdx.x * kScanItems;
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (first + k < nchunks) run += totals[first + k];
    double total = 0.0;
    block_exclusive_scan(run, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan (sum, or running maximum from 0 when is_max) of n values in
// place, by one block of kTotalsThreads: thread t takes a contiguous segment.
constexpr int kTotalsThreads = 1024;

__global__ void __launch_bounds__(kTotalsThreads) k_scan_small(double* __restrict__ v, uint64_t n, int is_max) {
    __shared__ double wsum[kTotalsThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t seg = (n + kTotalsThreads - 1) / kTotalsThreads;
    const uint64_t b0 = std::min<uint64_t>((uint64_t)threadIdx.x * seg, n), b1 = std::min<uint64_t>(b0 + seg, n);
    auto op = [is_max](double x, double y) { return is_max ? fmax(x, y) : x + y; };
    double acc = 0.0;
    for (uint64_t b = b0; b < b1; ++b) acc = op(acc, v[b]);
    double incl = acc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(incl, off, 64);
        if (lane >= off) incl = op(incl, y);
    }
    const double prev = __shfl_up(incl, 1, 64);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    double run = 0.0;
    for (int w = 0; w < wave; ++w) run = op(run, wsum[w]);
    if (lane > 0) run = op(run, prev);
    for (uint64_t b = b0; b < b1; ++b) {
        const double x = v[b];
        v[b] = run;
        run = op(run, x);
    }
}

// single: one group, whose offset and exclusive running maximum are 0 (the
// scans of k_totals_sums and k_scan_small would give exactly that); bmax[0]
// then receives that 0 directly.
__global__ void __launch_bounds__(kScanThreads) k_totals_apply(double* __restrict__ totals, double* __restrict__ tops,
                                                               uint64_t nchunks, const double* __restrict__ boff,
                                                               double* __restrict__ bmax, int single) {
    const uint64_t first = (uint64_t)blockIdx.x * kGroupChunks + (uint64_t)threadIdx.x * kScanItems;
    double sums[kScanItems];
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        sums[k] = first + k < nchunks ? totals[first + k] : 0.0;
        run += sums[k];
    }
    const double off = block_exclusive_scan(run, nullptr);
    const double group_off = single ? 0.0 : boff[blockIdx.x];
    double part = 0.0, mx = 0.0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const double o = (off + part) + group_off;
        part += sums[k];
        sums[k] = o;
        if (first + k < nchunks) mx = fmax(mx, tops[first + k] + o);
    }
    double all = 0.0;
    const double before = block_exclusive_max(mx, &all);
    double r = before;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (first + k < nchunks) {
            r = fmax(r, tops[first + k] + sums[k]);
            totals[first + k] = sums[k];
            tops[first + k] = r;
        }
    if (threadIdx.x == 0) bmax[blockIdx.x] = single ? 0.0 : all;
}

// Pass 3 (the written CDF): each chunk's prefixes plus its offset, raised
// to the running maximum of everything before them.
template <typename R>
__global__ void __launch_bounds__(kScanThreads) k_cdf_write(const cv<R>* __restrict__ psi, uint64_t count,
                                                            const double* __restrict__ offsets,
                                                            const double* __restrict__ tops,
                                                            const double* __restrict__ bpre, double* __restrict__ cdf) {
    __shared__ double sh[kScanChunk + kScanChunk / 16];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    double loc[kScanItems];
    const double off = chunk_prefixes<R>(psi, count, base, sh, loc, nullptr);
    const double chunk_off = offsets[blockIdx.x];
    const int nv = chunk_valid(count, base);
    double top = 0.0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        loc[k] = (off + loc[k]) + chunk_off;
        if (k < nv) top = loc[k];
    }
    const double floor = fmax(blockIdx.x > 0 ? chunk_top(tops, bpre, blockIdx.x - 1) : 0.0, block_exclusive_max(top, nullptr));
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sh[scan_lds(threadIdx.x * kScanItems + k)] = fmax(loc[k], floor);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t j = (uint32_t)k * kScanThreads + threadIdx.x;
        if (base + j < count) cdf[base + j] = sh[scan_lds(j)];
    }
}

// lower_bound (first index with cdf[idx] >= u) over the written CDF, clamped
// to count - 1 (reference core/src/statevector.cpp:248-254).
__global__ void __launch_bounds__(256) k_search(const double* __restrict__ cdf, uint64_t count, const double* __restrict__ u,
                                                uint64_t shots, uint64_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < shots; s += stride) {
        const double x = u[s];
        uint64_t lo = 0, hi = count;
        while (lo < hi) {
            const uint64_t mid = lo + ((hi - lo) >> 1);
            if (cdf[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        out[s] = lo < count ? lo : count - 1;
    }
}

// The same lower_bound without a written CDF, one block per shot: the first
// chunk whose running maximum (chunk_top) reaches u holds the answer; the block
// recomputes that chunk's prefixes exactly as k_cdf_write does and takes the
// first index whose value reaches u.  Reads one 4096-amplitude chunk per shot.
template <typename R>
__global__ void __launch_bounds__(kScanThreads) k_sample_chunks(const cv<R>* __restrict__ psi, uint64_t count,
                                                                const double* __restrict__ offsets,
                                                                const double* __restrict__ tops,
                                                                const double* __restrict__ bpre, uint64_t nchunks,
                                                                const double* __restrict__ u, uint64_t shots,
                                                                uint64_t* __restrict__ out) {
    __shared__ double sh[kScanChunk + kScanChunk / 16];
    __shared__ uint32_t first;
    for (uint64_t s = blockIdx.x; s < shots; s += gridDim.x) {
        const double x = u[s];
        uint64_t lo = 0, hi = nchunks;
        while (lo < hi) {
            const uint64_t mid = lo + ((hi - lo) >> 1);
            if (chunk_top(tops, bpre, mid) < x) lo = mid + 1;
            else hi = mid;
        }
        if (lo == nchunks) {  // above every CDF value: clamp (the same for the whole block)
            if (threadIdx.x == 0) out[s] = count - 1;
            continue;
        }
        const uint64_t base = lo * kScanChunk;
        if (threadIdx.x == 0) first = UINT32_MAX;
        double loc[kScanItems];
        const double off = chunk_prefixes<R>(psi, count, base, sh, loc, nullptr);  // barriers: `first` is set
        const double chunk_off = offsets[lo];
        const int nv = chunk_valid(count, base);
        int hit = -1;
#pragma unroll
        for (int k = kScanItems - 1; k >= 0; --k)
            if (k < nv && (off + loc[k]) + chunk_off >= x) hit = k;
        if (hit >= 0) atomicMin(&first, threadIdx.x * kScanItems + (uint32_t)hit);
        __syncthreads();
        if (threadIdx.x == 0) out[s] = first < kScanChunk ? base + first : count - 1;
        __syncthreads();  // sh and `first` are reused by the block's next shot
    }
}

// ---------------------------------------------------------- partial trace

struct KeepList {
    uint32_t q[16];
};

template <typename R>
__global__ void __launch_bounds__(256) k_partial_trace(const cv<R>* __restrict__ psi, KeepList keep, int k,
                                                       uint64_t envmask, double2* __restrict__ out) {
    const uint64_t kd = 1ULL << k;
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= kd * kd) return;
    const uint64_t r = id / kd, c = id % kd;
    uint64_t ir = 0, ic = 0;
    for (int b = 0; b < k; ++b) {
        if ((r >> b) & 1ULL) ir |= 1ULL << keep.q[b];
        if ((c >> b) & 1ULL) ic |= 1ULL << keep.q[b];
    }
    double sr = 0.0, si = 0.0;
    uint64_t e = 0;
    do {
        const cv<R> a = psi[ir | e];
        const cv<R> b = psi[ic | e];
        const double ar = a.x, ai = a.y, br = b.x, bi = b.y;
        sr += ar * br + ai * bi;  // a * conj(b)
        si += ai * br - ar * bi;
        e = (e - envmask) & envmask;
    } while (e != 0);
    out[id] = make_double2(sr, si);
}

// Amplitudes in (environment, kept) order, FP64: phi[e * 2^k + r] = psi[i]
// where bit b of r is qubit keep[b] of i and the bits of e are the traced
// qubits of i in ascending order (the index split of
// core/src/statevector.cpp:306-323).
template <typename R>
__global__ void __launch_bounds__(256) k_gather_keep(const cv<R>* __restrict__ psi, KeepList keep, int k,
                                                     uint64_t envmask, uint64_t total, double2* __restrict__ phi) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += stride) {
        const uint64_t r = o & ((1ULL << k) - 1), e = o >> k;
        uint64_t idx = deposit_bits(e, envmask);
        for (int b = 0; b < k; ++b) idx |= ((r >> b) & 1ULL) << keep.q[b];
        const cv<R> a = psi[idx];
        phi[o] = make_double2((double)a.x, (double)a.y);
    }
}

// rho_A = X^T conj(X) for X = phi viewed as m x d (row e, column r), on the
// FP64 matrix cores (v_mfma_f64_16x16x4): one wave per 16x16 tile of the upper
// triangle (tile row <= tile column), four e per step:
//   Re += Xr^T Xr + Xi^T Xi,  Im += Xi^T Xr - Xr^T Xi.
// Operand lanes: A holds row (lane % 16), k (lane / 16); B holds k (lane / 16),
// column (lane % 16); result value jj of a lane is row lane/16 + 4 jj, column
// lane % 16.  Off-diagonal tiles also write their conjugate transpose; a
// diagonal tile keeps its entries on or above the diagonal, each written with
// its conjugate transpose, and zero imaginary parts on the diagonal.
// Environment chunk blockIdx.z (rows [z * mchunk, (z + 1) * mchunk)) writes
// its partial matrix at out + z * slice.
__global__ void __launch_bounds__(64) k_herk_mfma(const double2* __restrict__ phi, uint64_t m, uint32_t d,
                                                  double2* __restrict__ out, uint64_t mchunk, uint64_t slice) {
    typedef double v4 __attribute__((ext_vector_type(4)));
    const uint32_t tr = blockIdx.y, tc = blockIdx.x;
    if (tc < tr) return;
    const uint64_t ebeg = (uint64_t)blockIdx.z * mchunk, eend = ebeg + mchunk < m ? ebeg + mchunk : m;
    out += (uint64_t)blockIdx.z * slice;
    const uint32_t lane = threadIdx.x & 63u, kk = lane >> 4, j = lane & 15u;
    const uint64_t r0 = (uint64_t)tr * 16 + j, c0 = (uint64_t)tc * 16 + j;
    v4 re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
#pragma unroll 4
    for (uint64_t e0 = ebeg; e0 < eend; e0 += 4) {
        const double2 a = phi[(e0 + kk) * d + r0];
        const double2 b = phi[(e0 + kk) * d + c0];
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, re, 0, 0, 0);
        re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, re, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, im, 0, 0, 0);
        im = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.x, b.y, im, 0, 0, 0);
    }
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const uint64_t row = (uint64_t)tr * 16 + kk + 4 * jj, col = (uint64_t)tc * 16 + j;
        if (row > col) continue;  // (diagonal tile) written by the transposed entry
        const double vi = row == col ? 0.0 : im[jj];
        out[row * d + col] = make_double2(re[jj], vi);
        if (row != col) out[col * d + row] = make_double2(re[jj], -vi);
    }
}

// The same product by 64x64 output blocks, for 6 or more kept and 4 or more
// traced qubits: a workgroup of four waves stages 16 environment rows of its
// row strip and its column strip in LDS per step (each row a 1 KiB
// contiguous load), and each wave accumulates a 32x32 quarter of the block
// as 2x2 tiles of v_mfma_f64_16x16x4 with the operand and result layouts of
// k_herk_mfma.  A strip element is read from memory once per block instead
// of once per 16x16 tile.  Off-diagonal blocks also write their conjugate
// transpose; in a diagonal block only the entries on or above the diagonal
// are kept, each written with its conjugate transpose, and the diagonal's
// imaginary parts (zero in exact arithmetic) as zero, so the matrix is
// exactly Hermitian.  Used below 64 blocks a side (fewer than 12 kept
// qubits); from there k_herk_dma takes over.
constexpr uint32_t kHerkBlock = 64, kHerkRows = 16;

__global__ void __launch_bounds__(256) k_herk_blocked(const double2* __restrict__ phi, uint64_t m, uint32_t d,
                                                      double2* __restrict__ out, uint64_t mchunk, uint64_t slice) {
    typedef double v4 __attribute__((ext_vector_type(4)));
    const uint32_t bR = blockIdx.y, bC = blockIdx.x;
    if (bC < bR) return;
    // environment chunk blockIdx.z: rows [z * mchunk, (z + 1) * mchunk), its
    // partial matrix at out + z * slice
    const uint64_t ebeg = (uint64_t)blockIdx.z * mchunk, eend = ebeg + mchunk < m ? ebeg + mchunk : m;
    out += (uint64_t)blockIdx.z * slice;
    __shared__ double2 sa[kHerkRows][kHerkBlock], sb[kHerkRows][kHerkBlock];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t wr = wave >> 1, wc = wave & 1u, kk = lane >> 4, j = lane & 15u;
    const uint64_t r0 = (uint64_t)bR * kHerkBlock, c0 = (uint64_t)bC * kHerkBlock;
    const bool diag = bR == bC;
    v4 re[2][2], im[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            re[a][b] = v4{0, 0, 0, 0};
            im[a][b] = v4{0, 0, 0, 0};
        }
    for (uint64_t e0 = ebeg; e0 < eend; e0 += kHerkRows) {
        // 16 rows x 64 of each strip: four elements per thread per strip
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t idx = tid + (uint32_t)i * 256u, er = idx >> 6, col = idx & 63u;
            const uint64_t row = (e0 + er) * d;
            sa[er][col] = phi[row + r0 + col];
            if (!diag) sb[er][col] = phi[row + c0 + col];
        }
        __syncthreads();
        const double2(*B)[kHerkBlock] = diag ? sa : sb;
#pragma unroll
        for (uint32_t ks = 0; ks < kHerkRows; ks += 4) {
            double2 a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = sa[ks + kk][wr * 32 + t * 16 + j];
                b[t] = B[ks + kk][wc * 32 + t * 16 + j];
            }
#pragma unroll
            for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                for (int tb = 0; tb < 2; ++tb) {
                    re[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].x, b[tb].x, re[ta][tb], 0, 0, 0);
                    re[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].y, b[tb].y, re[ta][tb], 0, 0, 0);
                    im[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].y, b[tb].x, im[ta][tb], 0, 0, 0);
                    im[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[ta].x, b[tb].y, im[ta][tb], 0, 0, 0);
                }
        }
        __syncthreads();  // the strips are rewritten by the next step
    }
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const uint64_t row = r0 + wr * 32 + ta * 16 + kk + 4 * jj, col = c0 + wc * 32 + tb * 16 + j;
                if (diag && row > col) continue;  // written by the transposed entry
                const double vr = re[ta][tb][jj], vi = row == col ? 0.0 : im[ta][tb][jj];
                out[row * d + col] = make_double2(vr, vi);
                if (row != col) out[col * d + row] = make_double2(vr, -vi);
            }
}

// The blocked product with its strips moved by the direct global-to-LDS
// loads of gfx950 (global_load_lds_dwordx4: each lane's 16 bytes land at
// the wave's LDS base + 16 * lane, no registers in between), double
// buffered: step s computes from one buffer while step s + 1's rows arrive
// in the other, one barrier per step.  The two buffers are distinct arrays
// used in a fixed alternation (and a diagonal block loads its strip into
// both), so the compiler can see that a step's LDS reads do not touch the
// loads in flight and issues no memory wait before them.  RW rows per step.
// Block order: workgroup b runs on XCD b % 8, and XCD x takes the
// super-tiles x, x + 8, ... of 8 x 8 blocks of the upper triangle, one at a
// time, so the blocks that run together on one XCD share 16 strips through
// its L2 (used from 8 super-tiles a side: with fewer, the work would sit on a
// few XCDs).  The strips come in under the products, so the kernel is bound
// by the matrix cores, and it takes three real products per tile where
// k_herk_blocked takes four:
//   re = Xr^T Xr,  im = Xi^T Xi,  t3 = (Xr + Xi)^T (Xr - Xi),
//   Re = re + im,  Im = Xi^T Xr - Xr^T Xi = t3 - re + im.
// 64 KiB of LDS per block at RW = 16; two blocks per CU are asked of the
// register allocator (226 VGPRs; unconstrained it kept two copies of the
// accumulators and fitted one wave per SIMD).  Environment chunks as
// k_herk_blocked.
template <int RW>
__global__ void __launch_bounds__(256, 2) k_herk_dma(const double2* __restrict__ phi, uint64_t m, uint32_t d,
                                                  double2* __restrict__ out, uint32_t ns, uint64_t mchunk,
                                                  uint64_t slice) {
    typedef double v4 __attribute__((ext_vector_type(4)));
    const uint32_t b = blockIdx.x, x = b & 7u, sq = b >> 3, w = sq & 63u;
    uint32_t g = (sq >> 6) * 8u + x, sr = 0;
    while (sr < ns && g >= ns - sr) {
        g -= ns - sr;
        ++sr;
    }
    if (sr >= ns) return;
    const uint32_t bR = sr * 8u + (w >> 3), bC = (sr + g) * 8u + (w & 7u);
    if (bC < bR) return;
    const uint64_t ebeg = (uint64_t)blockIdx.z * mchunk, eend = ebeg + mchunk < m ? ebeg + mchunk : m;
    out += (uint64_t)blockIdx.z * slice;
    // two buffers as distinct arrays, used in a fixed alternation, so that the
    // reads of one are not ordered behind the loads into the other
    __shared__ double2 sa0[RW][kHerkBlock], sb0[RW][kHerkBlock], sa1[RW][kHerkBlock], sb1[RW][kHerkBlock];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t wr = wave >> 1, wc = wave & 1u, kk = lane >> 4, j = lane & 15u;
    const uint64_t r0 = (uint64_t)bR * kHerkBlock, c0 = (uint64_t)bC * kHerkBlock;
    const bool diag = bR == bC;
    // rows of a step: wave w moves rows w, w + 4, ... of both strips, one
    // 1 KiB row (64 lanes x 16 B) per instruction
    auto fetch = [&](uint64_t e0, double2 (*SA)[kHerkBlock], double2 (*SB)[kHerkBlock]) __attribute__((always_inline)) {
#pragma unroll
        for (int er = (int)wave; er < RW; er += 4) {
            const double2* row = phi + (e0 + (uint64_t)er) * d;
            __builtin_amdgcn_global_load_lds((const void*)(row + r0 + lane),
                                             (__attribute__((address_space(3))) void*)&SA[er][0], 16, 0, 0);
            // (a diagonal block loads its strip twice: then every read has
            // one known buffer)
            __builtin_amdgcn_global_load_lds((const void*)(row + c0 + lane),
                                             (__attribute__((address_space(3))) void*)&SB[er][0], 16, 0, 0);
        }
    };
    v4 re[2][2], im[2][2], t3[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            re[a][b] = v4{0, 0, 0, 0};
            im[a][b] = v4{0, 0, 0, 0};
            t3[a][b] = v4{0, 0, 0, 0};
        }
    auto compute = [&](const double2 (*A)[kHerkBlock], const double2 (*B)[kHerkBlock]) __attribute__((always_inline)) {
#pragma unroll
        for (uint32_t ks = 0; ks < (uint32_t)RW; ks += 4) {
            double2 a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = A[ks + kk][wr * 32 + t * 16 + j];
                b[t] = B[ks + kk][wc * 32 + t * 16 + j];
            }
#pragma unroll
            for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                for (int tb = 0; tb < 2; ++tb) {
                    re[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].x, b[tb].x, re[ta][tb], 0, 0, 0);
                    im[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].y, b[tb].y, im[ta][tb], 0, 0, 0);
                    t3[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].x + a[ta].y, b[tb].x - b[tb].y, t3[ta][tb],
                                                                      0, 0, 0);
                }
        }
    };
    fetch(ebeg, sa0, sb0);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's rows are in LDS
    __syncthreads();
    // steps in pairs: buffer 0, then buffer 1 (the environment chunk is a
    // multiple of 16 rows, so a pair's second step exists when RW = 8, and for
    // RW = 16 it is checked)
    for (uint64_t e0 = ebeg; e0 < eend; e0 += 2 * RW) {
        const bool more1 = e0 + RW < eend, more2 = e0 + 2 * RW < eend;
        if (more1) fetch(e0 + RW, sa1, sb1);
        compute(sa0, sb0);
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (!more1) break;
        if (more2) fetch(e0 + 2 * RW, sa0, sb0);
        compute(sa1, sb1);
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
    }
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const uint64_t row = r0 + wr * 32 + ta * 16 + kk + 4 * jj, col = c0 + wc * 32 + tb * 16 + j;
                if (diag && row > col) continue;
                const double vr = re[ta][tb][jj] + im[ta][tb][jj];
                const double vi = row == col ? 0.0 : (t3[ta][tb][jj] - re[ta][tb][jj]) + im[ta][tb][jj];
                out[row * d + col] = make_double2(vr, vi);
                if (row != col) out[col * d + row] = make_double2(vr, -vi);
            }
}

// Real parts of the diagonal of a d x d matrix (double2, row-major) into out.
__global__ void __launch_bounds__(256) k_diag_real(const double2* __restrict__ A, uint32_t d, double* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d) out[i] = A[(size_t)i * d + i].x;
}

// Reduced matrices of 1-3 kept qubits: the environment sum split over the
// grid, chunk blockIdx.x taking environment indices [x * chunk, (x + 1) *
// chunk).  A lane takes one environment index at a time with its 2^K kept
// amplitudes in registers and adds the FP64 products of the upper triangle
// (entry u: row r, column c >= r, row by row); the lanes' sums are reduced
// in a fixed order (waves, then the workgroup) into part[x * U + u].
template <typename R, int K>
__global__ void __launch_bounds__(256) k_rho_small(const cv<R>* __restrict__ psi, KeepList keep, uint64_t envmask,
                                                   uint64_t m, uint64_t chunk, double2* __restrict__ part) {
    constexpr int D = 1 << K, U = D * (D + 1) / 2;
    __shared__ double s_acc[4][2 * U];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint64_t off[D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
        off[r] = 0;
#pragma unroll
        for (int b = 0; b < K; ++b)
            if ((r >> b) & 1) off[r] |= 1ULL << keep.q[b];
    }
    double acc[2 * U];
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) acc[u] = 0.0;
    const uint64_t e0 = (uint64_t)blockIdx.x * chunk, e1 = e0 + chunk < m ? e0 + chunk : m;
    for (uint64_t e = e0 + tid; e < e1; e += 256) {
        const uint64_t base = deposit_bits(e, envmask);
        double ar[D], ai[D];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            const cv<R> a = psi[base | off[r]];
            ar[r] = (double)a.x;
            ai[r] = (double)a.y;
        }
        int u = 0;
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = r; c < D; ++c, ++u) {
                acc[2 * u] += ar[r] * ar[c] + ai[r] * ai[c];  // a_r conj(a_c)
                acc[2 * u + 1] += ai[r] * ar[c] - ar[r] * ai[c];
            }
    }
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) {
        double v = acc[u];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) s_acc[wave][u] = v;
    }
    __syncthreads();
    if (tid < (uint32_t)(2 * U)) {
        const double v = ((s_acc[0][tid] + s_acc[1][tid]) + s_acc[2][tid]) + s_acc[3][tid];
        double* pd = reinterpret_cast<double*>(part + (uint64_t)blockIdx.x * U);
        pd[tid] = v;
    }
}

// The d x d matrix from `nparts` partial upper triangles (U = d (d + 1) / 2
// entries each, part[x * U + u]): one workgroup per entry, the partials summed
// in a fixed order; the entry and its conjugate transpose written, with a zero
// imaginary part on the diagonal.
__global__ void __launch_bounds__(256) k_rho_reduce(const double2* __restrict__ part, uint32_t nparts, uint32_t d,
                                                    double2* __restrict__ out) {
    __shared__ double s_re[4], s_im[4];
    const uint32_t u = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t U = d * (d + 1) / 2;
    double sr = 0.0, si = 0.0;
    for (uint32_t x = tid; x < nparts; x += 256) {
        const double2 v = part[(uint64_t)x * U + u];
        sr += v.x;
        si += v.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_down(sr, o, 64);
        si += __shfl_down(si, o, 64);
    }
    if (lane == 0) {
        s_re[wave] = sr;
        s_im[wave] = si;
    }
    __syncthreads();
    if (tid == 0) {
        sr = ((s_re[0] + s_re[1]) + s_re[2]) + s_re[3];
        si = ((s_im[0] + s_im[1]) + s_im[2]) + s_im[3];
        uint32_t r = 0, rest = u;
        while (rest >= d - r) {
            rest -= d - r;
            ++r;
        }
        const uint32_t c = r + rest;
        out[(uint64_t)r * d + c] = make_double2(sr, r == c ? 0.0 : si);
        if (r != c) out[(uint64_t)c * d + r] = make_double2(sr, -si);
    }
}

// out[i] = sum over s of part[s * count + i], s in order (the environment
// chunks of a split matrix-core product).
__global__ void __launch_bounds__(256) k_sum_slices(const double2* __restrict__ part, uint32_t nslices, uint64_t count,
                                                    double2* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double sr = 0.0, si = 0.0;
#pragma unroll 8
    for (uint32_t x = 0; x < nslices; ++x) {
        const double2 v = part[(uint64_t)x * count + i];
        sr += v.x;
        si += v.y;
    }
    out[i] = make_double2(sr, si);
}

// The capped Jacobi's rotation (jacobi_rows.hpp's backend, the loop of
// jacobi.hpp): one thread per row i != p, q.  The complex products are
// rounded as the host's complex arithmetic rounds them (each real product
// rounded, then added), with no fused multiply-add, so the device's matrix
// stays bit for bit the host's.
__global__ void __launch_bounds__(256) k_jacobi_rotate(double2* __restrict__ A, uint32_t d, uint32_t p, uint32_t q,
                                                       double c, double sr, double si, double npp, double nqq,
                                                       double2* __restrict__ colp, double2* __restrict__ colq,
                                                       uint32_t* __restrict__ extra_n, uint32_t ncol) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *extra_n = 0;
    if (i >= d) return;
    const size_t D = d;
    if (i == p) {
        A[(size_t)p * D + p] = make_double2(npp, 0.0);
        A[(size_t)p * D + q] = make_double2(0.0, 0.0);
        A[(size_t)q * D + p] = make_double2(0.0, 0.0);
        return;
    }
    if (i == q) {
        A[(size_t)q * D + q] = make_double2(nqq, 0.0);
        return;
    }
    const double2 aip = A[(size_t)i * D + p], aiq = A[(size_t)i * D + q];
    // c aip + conj(sn) aiq
    double2 np;
    np.x = c * aip.x + (sr * aiq.x - (-si) * aiq.y);
    np.y = c * aip.y + (sr * aiq.y + (-si) * aiq.x);
    // -sn aip + c aiq
    double2 nq;
    nq.x = ((-sr) * aip.x - (-si) * aip.y) + c * aiq.x;
    nq.y = ((-sr) * aip.y + (-si) * aip.x) + c * aiq.y;
    A[(size_t)i * D + p] = np;
    A[(size_t)p * D + i] = make_double2(np.x, -np.y);
    A[(size_t)i * D + q] = nq;
    A[(size_t)q * D + i] = make_double2(nq.x, -nq.y);
    if (i < ncol) {
        colp[i] = np;
        colq[i] = nq;
    }
}

// Candidates of the listed rows, one workgroup per row: the largest squared
// modulus right of the diagonal (top), then every entry at or above
// top * (1 - 1e-14) (every non-zero entry when top is below 1e-280), as
// jacobi_rows.hpp defines them.  More than kRowCands candidates: when they all
// share one unordered pair (|re|, |im|) (hence one std::abs) the row's answer
// is the first of them, else the host reads the row.
__global__ void __launch_bounds__(256) k_jacobi_scan(const double2* __restrict__ A, uint32_t d,
                                                     const uint32_t* __restrict__ rows, uint32_t nrows,
                                                     uint32_t* __restrict__ count, uint32_t* __restrict__ first_j,
                                                     double2* __restrict__ first_v, uint32_t* __restrict__ extra_n,
                                                     uint32_t cap, uint32_t* __restrict__ extra_t,
                                                     uint32_t* __restrict__ extra_j, double2* __restrict__ extra_v) {
#pragma clang fp contract(off)
    __shared__ double s_top[4];
    __shared__ uint32_t s_n, s_first, s_mixed, s_base;
    __shared__ uint32_t s_j[kRowCands];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (uint32_t t = blockIdx.x; t < nrows; t += gridDim.x) {
        const uint32_t i = rows[t];
        const double2* row = A + (size_t)i * d;
        double top = 0.0;
        for (uint32_t j = i + 1 + tid; j < d; j += 256) {
            const double2 z = row[j];
            top = fmax(top, z.x * z.x + z.y * z.y);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) top = fmax(top, __shfl_xor(top, off, 64));
        if (lane == 0) s_top[wave] = top;
        if (tid == 0) {
            s_n = 0;
            s_first = 0xffffffffu;
            s_mixed = 0;
        }
        __syncthreads();
        top = fmax(fmax(s_top[0], s_top[1]), fmax(s_top[2], s_top[3]));
        const double bound = top >= kJacobiFloor ? top * kJacobiSlack : -1.0;
        for (uint32_t j = i + 1 + tid; j < d; j += 256) {
            const double2 z = row[j];
            if (z.x * z.x + z.y * z.y >= bound && (z.x != 0.0 || z.y != 0.0)) {
                const uint32_t k = atomicAdd(&s_n, 1u);
                if (k < kRowCands) s_j[k] = j;
                atomicMin(&s_first, j);
            }
        }
        __syncthreads();
        const uint32_t n = s_n;
        if (n > kRowCands) {
            const double2 f = row[s_first];
            const double fa = fmax(fabs(f.x), fabs(f.y)), fb = fmin(fabs(f.x), fabs(f.y));
            for (uint32_t j = i + 1 + tid; j < d; j += 256) {
                const double2 z = row[j];
                if (z.x * z.x + z.y * z.y >= bound && (z.x != 0.0 || z.y != 0.0) &&
                    (fmax(fabs(z.x), fabs(z.y)) != fa || fmin(fabs(z.x), fabs(z.y)) != fb))
                    atomicOr(&s_mixed, 1u);
            }
            __syncthreads();
            if (tid == 0) {
                if (s_mixed) {
                    count[t] = kRowFetch;
                } else {
                    count[t] = 1;
                    first_j[t] = s_first;
                    first_v[t] = f;
                }
            }
        } else {
            if (tid == 0) {
                count[t] = n;
                if (n) {
                    first_j[t] = s_j[0];
                    first_v[t] = row[s_j[0]];
                }
                s_base = n > 1 ? atomicAdd(extra_n, n - 1) : 0u;
            }
            __syncthreads();
            if (tid >= 1 && tid < n) {
                const uint32_t k = s_base + tid - 1;
                if (k < cap) {
                    extra_t[k] = t;
                    extra_j[k] = s_j[tid];
                    extra_v[k] = row[s_j[tid]];
                }
            }
        }
        __syncthreads();  // the shared values are reset for the next row
    }
}

// Gate-matrix upload on the compute stream: reads pinned (coherent, uncached)
// host memory over PCIe and writes device memory, so the gate kernel that
// follows needs no copy-engine round trip.  One 16-byte element per thread
// per round trip: the grid covers the buffer in as few round trips as it can.
__global__ void __launch_bounds__(256) k_upload(const double2* __restrict__ src, double2* __restrict__ dst, uint32_t n2) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) dst[i] = src[i];
}

}  // namespace

// One launcher for the whole-state element-wise kernels: cached or streaming
// instantiation by state size, complex128 or complex64 by precision.
#define IQ_LAUNCH_PASS(KERNEL, count, prec, blocks, st, ptr_cast, ...)                                         \
    do {                                                                                                      \
        const bool nt_ = beyond_cache((count), (prec) == kC128 ? 16 : 8) && stream_hint_enabled();             \
        if ((prec) == kC128) {                                                                                \
            if (nt_) hipLaunchKernelGGL((KERNEL<double, true>), dim3(blocks), dim3(256), 0, st, (ptr_cast<double>)__VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<double, false>), dim3(blocks), dim3(256), 0, st, (ptr_cast<double>)__VA_ARGS__); \
        } else {                                                                                              \
            if (nt_) hipLaunchKernelGGL((KERNEL<float, true>), dim3(blocks), dim3(256), 0, st, (ptr_cast<float>)__VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<float, false>), dim3(blocks), dim3(256), 0, st, (ptr_cast<float>)__VA_ARGS__); \
        }                                                                                                     \
    } while (0)

template <typename R>
using amp_ptr = cv<R>*;
template <typename R>
using camp_ptr = const cv<R>*;

// element-wise passes: one thread per four amplitudes (every load of a thread
// in flight at once)
inline unsigned pass_blocks(uint64_t count) { return blocks_for((count + kPassSpan - 1) / kPassSpan, 256, 1ULL << 30); }

hipError_t launch_init_basis(void* psi, int n, int prec, uint64_t index, hipStream_t st) {
    const uint64_t count = 1ULL << n;
    const unsigned b = pass_blocks(count);
    IQ_LAUNCH_PASS(k_init_basis, count, prec, b, st, amp_ptr, psi, count, index);
    return hipGetLastError();
}

hipError_t launch_upload(void* dst, const void* src_host, size_t bytes, hipStream_t st) {
    if (bytes % 16 || bytes / 16 > 0xffffffffULL) return hipErrorInvalidValue;
    const uint32_t n2 = (uint32_t)(bytes / 16);
    const unsigned blocks = std::max(1u, std::min(1024u, (n2 + 255) / 256));
    hipLaunchKernelGGL(k_upload, dim3(blocks), dim3(256), 0, st, (const double2*)src_host, (double2*)dst, n2);
    return hipGetLastError();
}

hipError_t launch_collapse(void* psi, int n, int prec, int target, int outcome, double factor, hipStream_t st) {
    const uint64_t count = 1ULL << n;
    const unsigned b = pass_blocks(count);
    IQ_LAUNCH_PASS(k_collapse, count, prec, b, st, amp_ptr, psi, count, target, outcome, factor);
    return hipGetLastError();
}

hipError_t launch_scale(void* psi, uint64_t count, int prec, double factor, hipStream_t st) {
    const unsigned b = pass_blocks(count);
    IQ_LAUNCH_PASS(k_scale, count, prec, b, st, amp_ptr, psi, count, factor);
    return hipGetLastError();
}

static hipError_t split_strided(const void* psi, uint64_t first, uint64_t count, int prec, double* re, double* im,
                                uint32_t es, hipStream_t st) {
    const unsigned b = blocks_for(count, 256, 65536);
    if (prec == kC128)
        hipLaunchKernelGGL(k_split<double>, dim3(b), dim3(256), 0, st, (const double2*)psi + first, count, re, im, es);
    else hipLaunchKernelGGL(k_split<float>, dim3(b), dim3(256), 0, st, (const float2*)psi + first, count, re, im, es);
    return hipGetLastError();
}

static hipError_t join_strided(void* psi, uint64_t first, uint64_t count, int prec, const double* re, const double* im,
                               uint32_t es, hipStream_t st) {
    const unsigned b = blocks_for(count, 256, 65536);
    if (prec == kC128)
        hipLaunchKernelGGL(k_join<double>, dim3(b), dim3(256), 0, st, (double2*)psi + first, count, re, im, es);
    else hipLaunchKernelGGL(k_join<float>, dim3(b), dim3(256), 0, st, (float2*)psi + first, count, re, im, es);
    return hipGetLastError();
}

hipError_t launch_split(const void* psi, uint64_t first, uint64_t count, int prec, double* re, double* im, hipStream_t st) {
    return split_strided(psi, first, count, prec, re, im, 1, st);
}

hipError_t launch_join(void* psi, uint64_t first, uint64_t count, int prec, const double* re, const double* im,
                       hipStream_t st) {
    return join_strided(psi, first, count, prec, re, im, 1, st);
}

hipError_t launch_interleave(const void* psi, uint64_t first, uint64_t count, int prec, double* out, hipStream_t st) {
    return split_strided(psi, first, count, prec, out, out + 1, 2, st);
}

hipError_t launch_deinterleave(void* psi, uint64_t first, uint64_t count, int prec, const double* in, hipStream_t st) {
    return join_strided(psi, first, count, prec, in, in + 1, 2, st);
}

hipError_t launch_probs(const void* psi, uint64_t count, int prec, double* out, hipStream_t st) {
    const unsigned b = blocks_for(count, 256, 65536);
    if (prec == kC128) hipLaunchKernelGGL(k_probs<double>, dim3(b), dim3(256), 0, st, (const double2*)psi, count, out);
    else hipLaunchKernelGGL(k_probs<float>, dim3(b), dim3(256), 0, st, (const float2*)psi, count, out);
    return hipGetLastError();
}

static hipError_t finalize(const double* part, unsigned nblocks, int m, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, part, (int)nblocks, m, out);
    return hipGetLastError();
}

hipError_t launch_finalize(const double* part, unsigned nblocks, int m, double* out, hipStream_t st) {
    return finalize(part, nblocks, m, out, st);
}

hipError_t reduce_norm(const void* psi, uint64_t count, int prec, double* part, double* out, hipStream_t st) {
    const unsigned b = blocks_for(count, kRedThreads, kRedBlocks);
    static_assert(kRedThreads == 256, "the pass launcher uses 256 threads");
    IQ_LAUNCH_PASS(k_red_norm, count, prec, b, st, camp_ptr, psi, count, part);
    return finalize(part, b, 1, out, st);
}

hipError_t reduce_bit_probs(const void* psi, int n, int prec, int target, double* part, double* out, hipStream_t st) {
    const uint64_t count = 1ULL << n;
    const unsigned b = blocks_for(count, kRedThreads, kRedBlocks);
    IQ_LAUNCH_PASS(k_red_bitprob, count, prec, b, st, camp_ptr, psi, count, target, part);
    return finalize(part, b, 2, out, st);
}

hipError_t reduce_pauli_group(const void* psi, int n, int prec, uint64_t xmask, int nstr, const uint64_t* zy_masks,
                              const int* ny, double* part, double* out, hipStream_t st) {
    if (nstr < 1 || nstr > kPauliPerGroup) return hipErrorInvalidValue;
    PauliGroup g{};
    g.xmask = xmask;
    g.nstr = nstr;
    for (int s = 0; s < nstr; ++s) {
        g.zy[s] = zy_masks[s];
        g.ny[s] = ny[s];
    }
    const uint64_t count = 1ULL << n;
    const unsigned b = blocks_for(count, kRedThreads, kRedBlocks);
    // highest flipped bit: pairs (j, j ^ x) split on it
    const int hbit = xmask ? 63 - __builtin_clzll(xmask) : 0;
    // streaming when each load instruction covers whole 128-byte lines: the
    // pairs' low index runs are 2^hbit amplitudes long
    const int line_bits = prec == kC128 ? 3 : 4;
    const bool nt = beyond_cache(count, prec == kC128 ? 16 : 8) && stream_hint_enabled() && (xmask == 0 || hbit >= line_bits);
    if (prec == kC128) {
        if (nt)
            hipLaunchKernelGGL((k_red_pauli<double, true>), dim3(b), dim3(kRedThreads), 0, st, (const double2*)psi,
                               count, g, hbit, part);
        else
            hipLaunchKernelGGL((k_red_pauli<double, false>), dim3(b), dim3(kRedThreads), 0, st, (const double2*)psi,
                               count, g, hbit, part);
    } else if (nt) {
        hipLaunchKernelGGL((k_red_pauli<float, true>), dim3(b), dim3(kRedThreads), 0, st, (const float2*)psi, count,
                           g, hbit, part);
    } else {
        hipLaunchKernelGGL((k_red_pauli<float, false>), dim3(b), dim3(kRedThreads), 0, st, (const float2*)psi, count,
                           g, hbit, part);
    }
    return finalize(part, b, 2 * nstr, out, st);
}

hipError_t reduce_dm_diag(const void* rho, int nq, int mode, int target, double* part, double* out, hipStream_t st) {
    if (nq > 16) return hipErrorInvalidValue;
    const uint64_t dim = 1ULL << nq;
    const unsigned b = blocks_for(dim, kRedThreads, kRedBlocks);
    hipLaunchKernelGGL(k_red_dm_diag, dim3(b), dim3(kRedThreads), 0, st, (const double2*)rho, nq, mode, target, part);
    const int m = mode == 1 ? nq : (mode == 2 ? 2 : 1);
    return finalize(part, b, m, out, st);
}

hipError_t launch_dm_collapse(void* rho, int nq, int target, int outcome, double factor, hipStream_t st) {
    const uint64_t count = 1ULL << (2 * nq);
    const unsigned b = blocks_for(count, 256, 65536);
    hipLaunchKernelGGL(k_dm_collapse, dim3(b), dim3(256), 0, st, (double2*)rho, nq, target, outcome, factor);
    return hipGetLastError();
}

// Scan scratch: chunk offsets [nchunks], running maxima [nchunks], group
// offsets [ngroups], group running maxima [ngroups].
static uint64_t scan_chunks(uint64_t count) { return (count + kScanChunk - 1) / kScanChunk; }
static uint64_t scan_groups(uint64_t count) { return (scan_chunks(count) + kGroupChunks - 1) / kGroupChunks; }

struct ScanScratch {
    uint64_t nchunks, ngroups;
    double *offsets, *tops, *boff, *bpre;
    ScanScratch(uint64_t count, double* s)
        : nchunks(scan_chunks(count)), ngroups(scan_groups(count)), offsets(s), tops(s + nchunks),
          boff(s + 2 * nchunks), bpre(s + 2 * nchunks + ngroups) {}
};

size_t scan_scratch_doubles(uint64_t count) { return (size_t)(2 * scan_chunks(count) + 2 * scan_groups(count)); }

// A per-chunk search re-reads one chunk (4096 amplitudes); writing the CDF
// costs a read and an 8-byte write per amplitude.  Measured at 30 qubits
// complex128: 2^18 searches 2.5 ms, the CDF write 5.1 ms.
bool sample_without_cdf(uint64_t count, uint64_t shots) { return shots <= 2 * scan_chunks(count); }

template <typename R>
static hipError_t cdf_chunks(const void* psi, uint64_t count, const ScanScratch& z, hipStream_t st) {
    hipLaunchKernelGGL(k_cdf_sums<R>, dim3((unsigned)z.nchunks), dim3(kScanThreads), 0, st, (const cv<R>*)psi, count,
                       z.offsets, z.tops);
    if (z.ngroups == 1) {  // up to 4096 chunks (2^24 amplitudes): one launch
        hipLaunchKernelGGL(k_totals_apply, dim3(1), dim3(kScanThreads), 0, st, z.offsets, z.tops, z.nchunks, z.boff,
                           z.bpre, 1);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_totals_sums, dim3((unsigned)z.ngroups), dim3(kScanThreads), 0, st, z.offsets, z.nchunks, z.boff);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kTotalsThreads), 0, st, z.boff, z.ngroups, 0);
    hipLaunchKernelGGL(k_totals_apply, dim3((unsigned)z.ngroups), dim3(kScanThreads), 0, st, z.offsets, z.tops, z.nchunks,
                       z.boff, z.bpre, 0);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kTotalsThreads), 0, st, z.bpre, z.ngroups, 1);
    return hipGetLastError();
}

hipError_t launch_cdf_chunks(const void* psi, uint64_t count, int prec, double* scratch, hipStream_t st) {
    const ScanScratch z(count, scratch);
    if (z.nchunks > kMaxDispatchItems / kScanThreads) return hipErrorInvalidValue;  // one block per chunk
    return prec == kC128 ? cdf_chunks<double>(psi, count, z, st) : cdf_chunks<float>(psi, count, z, st);
}

hipError_t launch_cdf(const void* psi, uint64_t count, int prec, double* cdf, double* scratch, hipStream_t st) {
    const ScanScratch z(count, scratch);
    const hipError_t err = launch_cdf_chunks(psi, count, prec, scratch, st);
    if (err != hipSuccess) return err;
    if (prec == kC128)
        hipLaunchKernelGGL(k_cdf_write<double>, dim3((unsigned)z.nchunks), dim3(kScanThreads), 0, st, (const double2*)psi,
                           count, z.offsets, z.tops, z.bpre, cdf);
    else
        hipLaunchKernelGGL(k_cdf_write<float>, dim3((unsigned)z.nchunks), dim3(kScanThreads), 0, st, (const float2*)psi,
                           count, z.offsets, z.tops, z.bpre, cdf);
    return hipGetLastError();
}

hipError_t launch_search(const double* cdf, uint64_t count, const double* u, uint64_t shots, uint64_t* out, hipStream_t st) {
    if (shots == 0) return hipSuccess;
    const unsigned b = blocks_for(shots, 256, 1ULL << 30);
    hipLaunchKernelGGL(k_search, dim3(b), dim3(256), 0, st, cdf, count, u, shots, out);
    return hipGetLastError();
}

hipError_t launch_sample_chunks(const void* psi, uint64_t count, int prec, const double* scratch, const double* u,
                                uint64_t shots, uint64_t* out, hipStream_t st) {
    if (shots == 0) return hipSuccess;
    const ScanScratch z(count, const_cast<double*>(scratch));
    const unsigned b = (unsigned)std::min<uint64_t>(shots, 1ULL << 20);
    if (prec == kC128)
        hipLaunchKernelGGL(k_sample_chunks<double>, dim3(b), dim3(kScanThreads), 0, st, (const double2*)psi, count,
                           z.offsets, z.tops, z.bpre, z.nchunks, u, shots, out);
    else
        hipLaunchKernelGGL(k_sample_chunks<float>, dim3(b), dim3(kScanThreads), 0, st, (const float2*)psi, count,
                           z.offsets, z.tops, z.bpre, z.nchunks, u, shots, out);
    return hipGetLastError();
}

// Environment chunks for a product that would launch `blocks` workgroups over
// m environment rows: enough chunks for about `target` workgroups, at least
// 64 rows each (a multiple of 16), and no more partial slices of `slice`
// elements than `cap` holds.  Returns the chunk count (1: no split).
static uint32_t env_chunks(uint64_t blocks, uint64_t target, uint64_t m, uint64_t slice, uint64_t cap,
                           uint64_t& mchunk) {
    uint64_t S = blocks >= target ? 1 : (target + blocks - 1) / blocks;
    S = std::min<uint64_t>(S, std::max<uint64_t>(1, m / 64));
    S = std::min<uint64_t>(S, slice ? cap / slice : S);
    if (S <= 1) {
        mchunk = m;
        return 1;
    }
    mchunk = ((m + S - 1) / S + 15) / 16 * 16;
    return (uint32_t)((m + mchunk - 1) / mchunk);
}

hipError_t launch_partial_trace(const void* psi, int n, int prec, const uint32_t* keep, int k, double2* out,
                                double2* scratch, uint64_t scratch_elems, hipStream_t st) {
    if (k < 1 || k > 16 || k > n) return hipErrorInvalidValue;
    KeepList kl{};
    uint64_t kmask = 0;
    bool in_order = true;
    for (int b = 0; b < k; ++b) {
        kl.q[b] = keep[b];
        kmask |= 1ULL << keep[b];
        in_order &= keep[b] == (uint32_t)b;
    }
    const uint64_t envmask = ((1ULL << n) - 1) & ~kmask;
    const uint32_t d = 1u << k;
    const uint64_t m = 1ULL << (n - k);
    if (k <= 3 && m >= 2 && scratch) {
        // the environment sum over many workgroups, partial triangles in scratch
        const uint32_t U = d * (d + 1) / 2;
        uint64_t S = std::min<uint64_t>(2048, std::max<uint64_t>(1, m / 256));
        S = std::min<uint64_t>(S, std::max<uint64_t>(1, scratch_elems / U));
        const uint64_t chunk = (m + S - 1) / S;
        S = (m + chunk - 1) / chunk;
#define IQ_RHO_SMALL(RT, KK)                                                                                     \
    hipLaunchKernelGGL((k_rho_small<RT, KK>), dim3((unsigned)S), dim3(256), 0, st, (const cv<RT>*)psi, kl, envmask, m, \
                       chunk, scratch)
        if (prec == kC128) {
            if (k == 1) IQ_RHO_SMALL(double, 1);
            else if (k == 2) IQ_RHO_SMALL(double, 2);
            else IQ_RHO_SMALL(double, 3);
        } else {
            if (k == 1) IQ_RHO_SMALL(float, 1);
            else if (k == 2) IQ_RHO_SMALL(float, 2);
            else IQ_RHO_SMALL(float, 3);
        }
#undef IQ_RHO_SMALL
        hipLaunchKernelGGL(k_rho_reduce, dim3(U), dim3(256), 0, st, (const double2*)scratch, (uint32_t)S, d, out);
        return hipGetLastError();
    }
    const bool direct = in_order && prec == kC128;
    if (k >= 4 && n - k >= 2 && (direct || (scratch && scratch_elems >= (1ULL << n)))) {
        // matrix-core path: reorder to (environment, kept) unless already so
        const double2* phi = (const double2*)psi;
        double2* parts = scratch;  // environment-chunk partials after the reordered state
        uint64_t parts_cap = scratch ? scratch_elems : 0;
        if (!direct) {
            const uint64_t total = 1ULL << n;
            const unsigned b = blocks_for(total, 256, 65536);
            if (prec == kC128)
                hipLaunchKernelGGL(k_gather_keep<double>, dim3(b), dim3(256), 0, st, (const double2*)psi, kl, k, envmask,
                                   total, scratch);
            else
                hipLaunchKernelGGL(k_gather_keep<float>, dim3(b), dim3(256), 0, st, (const float2*)psi, kl, k, envmask,
                                   total, scratch);
            phi = scratch;
            parts = scratch + total;
            parts_cap = scratch_elems - total;
        }
        const uint64_t slice = (uint64_t)d * d;
        uint64_t mchunk = m;
        uint32_t S = 1;
        if (k >= 6 && n - k >= 4) {  // d a multiple of 64, m of 16
            const uint32_t nb = d / kHerkBlock;
            S = env_chunks((uint64_t)nb * (nb + 1) / 2, 512, m, slice, parts_cap, mchunk);
            double2* dst = S > 1 ? parts : out;
            const uint64_t sl = S > 1 ? slice : 0;
            // the XCD order from 8 super-tiles a side (every XCD gets several;
            // with fewer the work would sit on a few XCDs)
            if (nb >= 64) {
                // 12 or more kept: strips by direct global-to-LDS loads,
                // double buffered, three products per tile, in the XCD order
                // (15 of 30 kept: 1.81 against 2.69 s; 13 of 26: 44 against
                // 55 ms)
                const uint32_t ns = nb / 8, tiles = ns * (ns + 1) / 2;
                const uint32_t blocks = ((tiles + 7) / 8) * 8 * 64;
                hipLaunchKernelGGL((k_herk_dma<16>), dim3(blocks, 1, S), dim3(256), 0, st, phi, m, d, dst, ns, mchunk, sl);
            } else {
                hipLaunchKernelGGL(k_herk_blocked, dim3(nb, nb, S), dim3(256), 0, st, phi, m, d, dst, mchunk, sl);
            }
        } else {
            const uint32_t nt = d / 16;
            // one wave per tile: chunks for about 2048 waves
            S = env_chunks((uint64_t)nt * (nt + 1) / 2, 2048, m, slice, parts_cap, mchunk);
            double2* dst = S > 1 ? parts : out;
            hipLaunchKernelGGL(k_herk_mfma, dim3(nt, nt, S), dim3(64), 0, st, phi, m, d, dst, mchunk, S > 1 ? slice : 0);
        }
        if (S > 1)
            hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((slice + 255) / 256)), dim3(256), 0, st, (const double2*)parts,
                               S, slice, out);
        return hipGetLastError();
    }
    const uint64_t total = (uint64_t)d * d;
    const unsigned b = blocks_for(total, 256, 1ULL << 30);
    if (prec == kC128)
        hipLaunchKernelGGL(k_partial_trace<double>, dim3(b), dim3(256), 0, st, (const double2*)psi, kl, k, envmask, out);
    else hipLaunchKernelGGL(k_partial_trace<float>, dim3(b), dim3(256), 0, st, (const float2*)psi, kl, k, envmask, out);
    return hipGetLastError();
}

hipError_t launch_jacobi_rotate(double2* A, uint32_t d, uint32_t p, uint32_t q, double c, double sr, double si,
                                double npp, double nqq, double2* colp, double2* colq, uint32_t* extra_n, uint32_t ncol,
                                hipStream_t st) {
    if (d < 2 || p >= d || q >= d || p == q || ncol > d) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_jacobi_rotate, dim3((d + 255) / 256), dim3(256), 0, st, A, d, p, q, c, sr, si, npp, nqq, colp,
                       colq, extra_n, ncol);
    return hipGetLastError();
}

hipError_t launch_jacobi_scan(const double2* A, uint32_t d, const uint32_t* rows, uint32_t nrows, uint32_t* count,
                              uint32_t* first_j, double2* first_v, uint32_t* extra_n, uint32_t cap, uint32_t* extra_t,
                              uint32_t* extra_j, double2* extra_v, hipStream_t st) {
    if (nrows == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>(nrows, 16384u);
    hipLaunchKernelGGL(k_jacobi_scan, dim3(blocks), dim3(256), 0, st, A, d, rows, nrows, count, first_j, first_v,
                       extra_n, cap, extra_t, extra_j, extra_v);
    return hipGetLastError();
}

hipError_t launch_diag_real(const double2* A, uint32_t d, double* out, hipStream_t st) {
    if (d == 0) return hipSuccess;
    hipLaunchKernelGGL(k_diag_real, dim3((d + 255) / 256), dim3(256), 0, st, A, d, out);
    return hipGetLastError();
}

}  // namespace iqg
