// Gate application kernels for gfx950: the HBM-bound pass over 2^n amplitudes.
//
// A K-qubit gate (K = 1..5) splits the state into 2^(n-K) independent groups of
// 2^K amplitudes; each group is multiplied by the 2^K x 2^K matrix.  This is
// the algorithm of reference core/src/statevector.cpp:26-152 (1q pairs, 2q
// quads, fused gather/matvec/scatter) with the documented MSB-first local
// index for every K: local bit (K-1-m) of a group member is qubit q[m].
//
// Two kernels:
//   direct  one lane = one group.  The group is loaded straight into VGPRs.
//           Lanes of a wave span the lowest non-target qubits, so each load
//           instruction is coalesced when the lowest target is >= 3 (c128) /
//           >= 4 (c64), i.e. runs of >= 128 B.
//   tiled   one wave stages a 2^(K+6)-amplitude tile in LDS: the targets plus
//           the six lowest non-target qubits.  Those six always include qubits
//           0..5, so every global access is a contiguous run of >= 64
//           amplitudes.  Lanes then gather their group from LDS (XOR swizzle
//           keeps the 16-lane phases bank-conflict free), multiply, write back
//           to LDS and the wave stores the tile coalesced.
// The matrix is read with uniform (scalar) loads, so every FMA takes its
// matrix operand from an SGPR pair.  K >= 4 blocks use the 3-multiply complex
// product (Re = A.u - B.v, Im = (A+B).(u+v) - A.u - B.v): 3 real matvecs
// instead of 4, which moves the FP64 compute roof from 9.8 to ~13 TB/s.

#include "kernels.hpp"
#include "layout.hpp"
#include "memio.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace iqg {

namespace {

template <typename R, int N>
struct ValMat {
    R v[N];
};

// XCD-consistent split of a launch's work items, for states that fit the
// eight XCDs' L2 caches together.  The dispatcher hands workgroup b to XCD
// b % 8; with xm != 0 the grid is exactly 2^xm items in a multiple of 8
// blocks, and block b takes the (b % 8)-th eighth of the items.  The top three
// bits of an item index are the top three bits of its amplitudes' indices
// whenever no target sits there, so every gate kernel hands the amplitudes
// with top bits x to XCD x, whose L2 still holds them from the last gate.
// xm = 0: item blockIdx * blockDim + threadIdx, grid-stride.
__device__ __forceinline__ uint64_t xcd_item(uint32_t xm) {
    if (!xm) return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    return ((uint64_t)(blockIdx.x & 7u) << (xm - 3)) | ((uint64_t)(blockIdx.x >> 3) * blockDim.x + threadIdx.x);
}

// y = U x for one group held in registers.  Rows run in a runtime loop for
// D >= 8: each iteration streams one matrix row through SGPRs (scalar loads),
// which bounds SGPR pressure; a fully unrolled 32x32 product would hoist all
// 2048-3072 coefficients and spill.  Two partial sums per chain shorten the
// dependent-FMA chains.
template <typename R, int D, bool M3, typename Get, typename Put>
__device__ __forceinline__ void matvec(const R (&xr)[D], const R (&xi)[D], Get m, Put put) {
    constexpr int kW = M3 ? 3 * D : 2 * D;  // coefficients per row
    if constexpr (M3) {
        R xs[D];
#pragma unroll
        for (int c = 0; c < D; ++c) xs[c] = xr[c] + xi[c];
#pragma unroll(D >= 8 ? 1 : D)
        for (int r = 0; r < D; ++r) {
            R t1a = 0, t2a = 0, t3a = 0, t1b = 0, t2b = 0, t3b = 0;
#pragma unroll
            for (int c = 0; c < D; c += 2) {
                t1a = fma(m(r * kW + c), xr[c], t1a);
                t2a = fma(m(r * kW + D + c), xi[c], t2a);
                t3a = fma(m(r * kW + 2 * D + c), xs[c], t3a);
                if (c + 1 < D) {
                    t1b = fma(m(r * kW + c + 1), xr[c + 1], t1b);
                    t2b = fma(m(r * kW + D + c + 1), xi[c + 1], t2b);
                    t3b = fma(m(r * kW + 2 * D + c + 1), xs[c + 1], t3b);
                }
            }
            const R t1 = t1a + t1b, t2 = t2a + t2b, t3 = t3a + t3b;
            put(r, t1 - t2, t3 - t1 - t2);
        }
    } else {
#pragma unroll(D >= 8 ? 1 : D)
        for (int r = 0; r < D; ++r) {
            R yra = 0, yia = 0, yrb = 0, yib = 0;
#pragma unroll
            for (int c = 0; c < D; c += 2) {
                const R a = m(r * kW + c);
                const R b = m(r * kW + D + c);
                yra = fma(a, xr[c], yra);
                yra = fma(-b, xi[c], yra);
                yia = fma(a, xi[c], yia);
                yia = fma(b, xr[c], yia);
                if (c + 1 < D) {
                    const R a1 = m(r * kW + c + 1);
                    const R b1 = m(r * kW + D + c + 1);
                    yrb = fma(a1, xr[c + 1], yrb);
                    yrb = fma(-b1, xi[c + 1], yrb);
                    yib = fma(a1, xi[c + 1], yib);
                    yib = fma(b1, xr[c + 1], yib);
                }
            }
            put(r, yra + yrb, yia + yib);
        }
    }
}

// One- and two-qubit gates whose targets all lie below bit 6, inside the 64
// consecutive amplitudes of a wave: lane l loads amplitude base + l, so each
// wave moves one contiguous run per instruction (the direct kernel's lanes
// read strided parts of lines when a target is inside a line, and two or four
// runs otherwise), takes its group's other members from the lanes holding
// them, and stores its own output row.  The product is matvec's on the same values, so the result
// equals the direct kernel's bit for bit.
template <int K>
struct LaneArgs {
    uint32_t q[K];           // targets, list order (local bit K-1-m <-> q[m])
    uint32_t tmask;          // lane bits of the targets
    uint32_t offl[1 << K];   // lane offset of local index c
};

template <typename R, int K, bool NT>
__global__ void __launch_bounds__(256) k_gate_lanes(cv<R>* __restrict__ psi, uint64_t count, LaneArgs<K> a,
                                                    ValMat<R, 2 * (1 << (2 * K))> mat) {
    constexpr int D = 1 << K;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t l = 0;  // this lane's local index
#pragma unroll
    for (int m = 0; m < K; ++m) l |= ((lane >> a.q[m]) & 1u) << (K - 1 - m);
    const uint32_t others = lane & ~a.tmask;
    // count is a multiple of 64 and every wave starts 64-aligned, so the
    // lanes of a wave run the loop together (the shuffles need all of them)
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const cv<R> v = ld_stream<R>(psi + i, NT);
        R xr[D], xi[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const int src = (int)(others | a.offl[c]);
            xr[c] = __shfl(v.x, src, 64);
            xi[c] = __shfl(v.y, src, 64);
        }
        cv<R> o;
        o.x = v.x;
        o.y = v.y;
        matvec<R, D, false>(
            xr, xi, [&](int j) { return mat.v[j]; },
            [&](int r, R yr, R yi) {
                if ((uint32_t)r == l) {
                    o.x = yr;
                    o.y = yi;
                }
            });
        st_stream<R>(psi + i, o, NT);
    }
}

// Two-qubit operations with one target tl below bit 6 and the other, th, at
// bit 6 or above (a density matrix's one-qubit channels on qubits 0-5:
// superoperators on qubits t and t + n): lane l holds the two amplitudes
// i0 = base + l and i0 + 2^th (two runs of whole lines, where the direct
// kernel's four group members make four), takes the tl-partners of both from
// lane l ^ 2^tl, and stores the two rows whose tl bit is its own.  matvec's
// product on the same values: the direct kernel's result bit for bit.
template <typename R, bool NT>
__global__ void __launch_bounds__(256) k_gate_lanes_mixed(cv<R>* __restrict__ psi, uint64_t npairs, uint32_t tl,
                                                          uint32_t th, int low_first, ValMat<R, 32> mat) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t lb = (lane >> tl) & 1u;
    const int src = (int)(lane ^ (1u << tl));
    const uint64_t lo = (1ULL << th) - 1;
    // npairs = 2^(n-1) is a multiple of 64 and waves start 64-aligned: the
    // lanes of a wave run the loop together
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < npairs; g += stride) {
        const uint64_t i0 = ((g & ~lo) << 1) | (g & lo), i1 = i0 | (1ULL << th);
        const cv<R> v0 = ld_stream<R>(psi + i0, NT), v1 = ld_stream<R>(psi + i1, NT);
        cv<R> p0, p1;
        p0.x = __shfl(v0.x, src, 64);
        p0.y = __shfl(v0.y, src, 64);
        p1.x = __shfl(v1.x, src, 64);
        p1.y = __shfl(v1.y, src, 64);
        // x[c], c the MSB-first local index over the targets in list order
        R xr[4], xi[4];
#pragma unroll
        for (int low = 0; low < 2; ++low)
#pragma unroll
            for (int high = 0; high < 2; ++high) {
                const cv<R>& mine = high ? v1 : v0;
                const cv<R>& other = high ? p1 : p0;
                const bool own = (uint32_t)low == lb;
                const int c = low_first ? (low << 1) | high : (high << 1) | low;
                xr[c] = own ? mine.x : other.x;
                xi[c] = own ? mine.y : other.y;
            }
        cv<R> o0 = v0, o1 = v1;
        matvec<R, 4, false>(
            xr, xi, [&](int j) { return mat.v[j]; },
            [&](int r, R yr, R yi) {
                const uint32_t low = low_first ? ((uint32_t)r >> 1) : ((uint32_t)r & 1u);
                const uint32_t high = low_first ? ((uint32_t)r & 1u) : ((uint32_t)r >> 1);
                if (low == lb) {
                    if (high) {
                        o1.x = yr;
                        o1.y = yi;
                    } else {
                        o0.x = yr;
                        o0.y = yi;
                    }
                }
            });
        st_stream<R>(psi + i0, o0, NT);
        st_stream<R>(psi + i1, o1, NT);
    }
}

// NT: streaming loads and stores, for target sets with no target inside a
// 128-byte line (each load instruction then covers whole lines).
template <typename R, int K, bool M3, bool NT = false>
__global__ void __launch_bounds__(256) k_gate_direct_ptr(cv<R>* __restrict__ psi, uint64_t ngroups, DirectArgs<K> a,
                                                         const R* __restrict__ mat, uint32_t xm) {
    constexpr int D = 1 << K;
    // one group per thread; the grid is capped above 2^32 threads
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = xcd_item(xm); g < ngroups; g += stride) {
        const uint64_t base = insert_zero_bits<K>(g, a.pos);
        R xr[D], xi[D];
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const cv<R> v = ld_stream<R>(psi + base + a.off[s], NT);
            xr[s] = v.x;
            xi[s] = v.y;
        }
        matvec<R, D, M3>(
            xr, xi, [&](int i) { return mat[i]; },
            [&](int r, R yr, R yi) {
                cv<R> o;
                o.x = yr;
                o.y = yi;
                st_stream<R>(psi + base + a.off[r], o, NT);
            });
    }
}

// NT as in k_gate_direct_ptr.
template <typename R, int K, bool M3, bool NT = false>
__global__ void __launch_bounds__(256) k_gate_direct_val(cv<R>* __restrict__ psi, uint64_t ngroups, DirectArgs<K> a,
                                                         ValMat<R, (M3 ? 3 : 2) * (1 << (2 * K))> mat, uint32_t xm) {
    constexpr int D = 1 << K;
    // one group per thread; the grid is capped above 2^32 threads
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = xcd_item(xm); g < ngroups; g += stride) {
        const uint64_t base = insert_zero_bits<K>(g, a.pos);
        R xr[D], xi[D];
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const cv<R> v = ld_stream<R>(psi + base + a.off[s], NT);
            xr[s] = v.x;
            xi[s] = v.y;
        }
        matvec<R, D, M3>(
            xr, xi, [&](int i) { return mat.v[i]; },
            [&](int r, R yr, R yi) {
                cv<R> o;
                o.x = yr;
                o.y = yi;
                st_stream<R>(psi + base + a.off[r], o, NT);
            });
    }
}

template <typename R, int K, bool M3>
__global__ void __launch_bounds__(64) k_gate_tiled(cv<R>* __restrict__ psi, TiledArgs<K> a, const R* __restrict__ mat,
                                                   uint64_t t0) {
    constexpr int D = 1 << K;
    constexpr int T = 64 * D;
    __shared__ cv<R> tile[T];
    const uint32_t lane = threadIdx.x;
    uint32_t pl, pg;
    tile_lane_positions<K>(lane, a, pl, pg);
    const uint64_t t = t0 + blockIdx.x;  // one tile per block
    const uint64_t base = deposit_bits(t, a.rest) + lane;
    {
        cv<R> v[D];
#pragma unroll
        for (int m = 0; m < D; ++m) v[m] = psi[base + a.off[m]];
#pragma unroll
        for (int m = 0; m < D; ++m) tile[pl ^ a.pm[m]] = v[m];
    }
    __syncthreads();
    R xr[D], xi[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        const cv<R> x = tile[pg ^ a.ts[s]];
        xr[s] = x.x;
        xi[s] = x.y;
    }
    matvec<R, D, M3>(
        xr, xi, [&](int i) { return mat[i]; },
        [&](int r, R yr, R yi) {
            cv<R> o;
            o.x = yr;
            o.y = yi;
            tile[pg ^ a.ts[r]] = o;
        });
    __syncthreads();
#pragma unroll
    for (int m = 0; m < D; ++m) psi[base + a.off[m]] = tile[pl ^ a.pm[m]];
}

template <typename R>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// 5-qubit block on the matrix cores, 3-multiply form: P1 = A Xr, P2 = B Xi,
// P3 = (A+B)(Xr+Xi) as 32x32 by 32x16 real products (2 row tiles x 8 k-steps
// of 16x16x4 each, 48 MFMAs per 16 groups); Re y = P1 - P2, Im y = P3 - P1 - P2.
// The matrix fragments (48 per lane) stay in VGPRs for the whole kernel, so
// the VALU only does addressing and the final combination, and memory traffic
// of one batch overlaps the matrix-core work of another wave.
template <typename R, int UB, bool WLDS, bool NT>
__global__ void __launch_bounds__(256) k_gate_mfma(cv<R>* __restrict__ psi, MfmaArgs a, const R* __restrict__ frag) {
    using M = Mfma<R>;
    using v4 = typename M::v4;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t q = lane >> 4, nn = lane & 15u;
    // Matrix fragments: 48 per lane, in VGPRs or shared by the block's waves in LDS.
    __shared__ R wsh[WLDS ? 48 * 64 : 1];
    R wreg[WLDS ? 1 : 48];
    if constexpr (WLDS) {
        for (uint32_t i = threadIdx.x; i < 48u * 64u; i += blockDim.x) wsh[i] = frag[i];
        __syncthreads();
    } else {
#pragma unroll
        for (int f = 0; f < 48; ++f) wreg[f] = frag[f * 64 + lane];
    }
    uint32_t wl = lane;
    auto W = [&](int f) -> R {
        if constexpr (WLDS) return wsh[f * 64 + wl];
        else return wreg[f];
    };
    const uint64_t xq = a.xq[q], yq = a.yq[q];
    // with a.xm, one super-batch per wave and block b takes the (b % 8)-th
    // eighth of them (xcd_item's split: the super-batch's top bits are its
    // amplitudes' top bits)
    const uint64_t wave = a.xm ? (((uint64_t)(blockIdx.x & 7u) << (a.xm - 3)) |
                                  ((uint64_t)(blockIdx.x >> 3) * (blockDim.x >> 6) + (threadIdx.x >> 6)))
                               : ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t nsuper = a.nbatches / UB;
    for (uint64_t sb0 = wave; sb0 < nsuper; sb0 += nwaves) {
        const uint64_t sb = rotate_index(sb0, a.rot, a.bits);
        // An opaque copy of the lane index per iteration keeps the LDS fragment
        // reads inside the loop (otherwise they are hoisted into 96 VGPRs).
        if constexpr (WLDS) asm volatile("" : "+v"(wl));
        uint64_t base[UB];
        R xr[UB][8], xi[UB][8];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            base[u] = insert_zero_bits<5>((sb * UB + u) * 16 + nn, a.pos);
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) {
                const cv<R> v = ld_stream<R>(psi + base[u] + xq + a.xk[kb], NT);
                xr[u][kb] = v.x;
                xi[u][kb] = v.y;
            }
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            v4 p1[2], p2[2], p3[2];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                p1[mt] = v4{0, 0, 0, 0};
                p2[mt] = v4{0, 0, 0, 0};
                p3[mt] = v4{0, 0, 0, 0};
            }
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) {
                const R xs = xr[u][kb] + xi[u][kb];
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) {
                    p1[mt] = M::op(W((0 * 2 + mt) * 8 + kb), xr[u][kb], p1[mt]);
                    p2[mt] = M::op(W((1 * 2 + mt) * 8 + kb), xi[u][kb], p2[mt]);
                    p3[mt] = M::op(W((2 * 2 + mt) * 8 + kb), xs, p3[mt]);
                }
            }
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    cv<R> o;
                    o.x = p1[mt][jj] - p2[mt][jj];
                    o.y = p3[mt][jj] - p1[mt][jj] - p2[mt][jj];
                    st_stream<R>(psi + base[u] + yq + a.yk[mt * 4 + jj], o, NT);
                }
        }
    }
}

// Controlled one-qubit gate, a K-qubit matrix blockdiag(I, U) over K-1
// control bits: the identity block leaves its amplitudes unchanged, so only
// the 2^-(K-1) of the state with every control bit set is read and written
// (pairs base, base + toff with the control bits set and the target bit clear).
template <typename R, int K, bool NT>
__global__ void __launch_bounds__(256) k_gate_controlled(cv<R>* __restrict__ psi, uint64_t npairs, DirectArgs<K> a,
                                                         uint64_t cbits, uint64_t toff, ValMat<R, 8> mat, uint32_t xm) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = xcd_item(xm); g < npairs; g += stride) {
        const uint64_t base = insert_zero_bits<K>(g, a.pos) | cbits;
        R xr[2], xi[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const cv<R> v = ld_stream<R>(psi + base + (s ? toff : 0), NT);
            xr[s] = v.x;
            xi[s] = v.y;
        }
        matvec<R, 2, false>(
            xr, xi, [&](int i) { return mat.v[i]; },
            [&](int r, R yr, R yi) {
                cv<R> o;
                o.x = yr;
                o.y = yi;
                st_stream<R>(psi + base + (r ? toff : 0), o, NT);
            });
    }
}

// Diagonal operator: each amplitude times the entry its target bits select.
// Four amplitudes in flight per thread; NT streams them (large states: each
// amplitude is touched once).
template <typename R, int K, bool NT = false>
__global__ void __launch_bounds__(256) k_diag(cv<R>* __restrict__ psi, uint64_t count, DirectArgs<K> a,
                                              ValMat<double, 2 * (1 << K)> d, uint32_t xm) {
    constexpr int D = 1 << K;
    __shared__ double sre[D], sim[D];
    if (threadIdx.x < D) {
        sre[threadIdx.x] = d.v[threadIdx.x];
        sim[threadIdx.x] = d.v[D + threadIdx.x];
    }
    __syncthreads();
    auto one = [&](uint64_t i, cv<R> v) __attribute__((always_inline)) {
        uint32_t loc = 0;
#pragma unroll
        for (int m = 0; m < K; ++m) loc |= (uint32_t)((i >> a.pos[m]) & 1u) << (K - 1 - m);
        const R dr = (R)sre[loc], di = (R)sim[loc];
        cv<R> o;
        o.x = dr * v.x - di * v.y;
        o.y = dr * v.y + di * v.x;
        st_stream<R>(psi + i, o, NT);
    };
    if (xm) {  // one amplitude per thread, XCD-consistent
        const uint64_t i = xcd_item(xm);
        one(i, ld_stream<R>(psi + i, NT));
        return;
    }
    // a block takes spans of kPassSpan * blockDim consecutive amplitudes, so a
    // wave's accesses fall in one span
    const uint64_t span = (uint64_t)kPassSpan * blockDim.x;
    for (uint64_t b = (uint64_t)blockIdx.x * span; b < count; b += (uint64_t)gridDim.x * span) {
        const uint64_t i = b + threadIdx.x;
        if (b + span <= count) {
            cv<R> v[kPassSpan];
#pragma unroll
            for (int u = 0; u < kPassSpan; ++u) v[u] = ld_stream<R>(psi + i + u * blockDim.x, NT);
#pragma unroll
            for (int u = 0; u < kPassSpan; ++u) one(i + u * blockDim.x, v[u]);
        } else {
            for (uint64_t j = i; j < count; j += blockDim.x) one(j, ld_stream<R>(psi + j, NT));
        }
    }
}

// blocks for `items` threads, capped below 2^32 threads (the kernels loop)
inline unsigned grid_for(uint64_t items, unsigned block) {
    const uint64_t g = std::min<uint64_t>((items + block - 1) / block, kMaxDispatchItems / block);
    return (unsigned)std::max<uint64_t>(1, g);
}

static int g_direct_nt = 1;

// States up to this size (half the eight XCDs' L2 caches together) take the
// XCD-consistent work split (xcd_item).  Measured on a 20-qubit complex128
// state (16 MiB), one-qubit gates on random targets below the top three
// qubits: 5.3 against 8.1 us per gate.
constexpr uint64_t kXcdStateMax = 16ULL << 20;
constexpr unsigned kXcds = 8;  // XCDs of the MI355X (gfx950)

// xm for a launch of `items` work items (one per thread) in `blocks` blocks of
// 256, over a state of `state_bytes`: log2(items) when the split applies (a
// power-of-two item count, exactly one item per thread, a multiple of eight
// blocks), else 0.
inline uint32_t xcd_log2(uint64_t items, unsigned blocks, uint64_t state_bytes) {
    if (state_bytes > kXcdStateMax || (uint64_t)blocks * 256 != items || blocks % kXcds != 0 || (items & (items - 1)))
        return 0;
    uint32_t m = 0;
    while ((1ULL << m) < items) ++m;
    return m;
}  // streaming hint in the direct kernel: 1 where lines are whole, 0 never
constexpr uint32_t kLaneLow = 6;  // lane-exchange forms for targets below this bit (a wave's 64 lanes)

template <typename R, int K, bool M3>
hipError_t run_direct(void* psi, int n, hipStream_t st, const GateDesc& g) {
    DirectArgs<K> a;
    fill_direct_args<K>(a, g.q);
    const uint64_t ngroups = 1ULL << (n - K);
    const unsigned blocks = grid_for(ngroups, 256);
    constexpr int D = 1 << K;
    const uint32_t xm = xcd_log2(ngroups, blocks, (1ULL << n) * sizeof(cv<R>));
    // streaming when no target sits inside a 128-byte line and the state is
    // beyond the 256 MiB Infinity Cache (measured on a 14-qubit density
    // matrix, 2^28 complex128 entries: one-qubit Kraus channels 5.43 against
    // 5.27 TB/s, one-qubit unitaries 5.65 against 5.45).  A state that fits
    // stays there between gates when cached: per random one- or two-qubit
    // gate, complex128 24 qubits 63-67 us against 79-87 with the hints,
    // complex64 25 qubits 66-73 against 82-93; at 25 qubits complex128 the
    // hints win, 159-176 against 177-191 us.
    const uint32_t line_bits = std::is_same<R, double>::value ? 3u : 4u;
    const bool stream = ((1ULL << n) * sizeof(cv<R>)) > kCacheStateMax && g_direct_nt != 0;
    bool whole = stream;
    for (int m = 0; m < K; ++m) whole = whole && g.q[m] >= line_bits;
    if constexpr (K <= 2) {
        // Lane-exchange forms for targets below bit 6, on states beyond the
        // 256 MiB Infinity Cache (measured at 14-qubit density matrices and 30
        // qubits).  On states that stay in the cache the shuffles' latency
        // made them slower: 20 qubits 12.3 against 7.9 us per pass, complex64
        // 25 qubits 71.7 against 65.5 us per one-qubit gate.
        const bool big = beyond_cache(1ULL << n, sizeof(cv<R>));
        bool lanes = big;  // every target below bit 6
        for (int m = 0; m < K; ++m) lanes = lanes && g.q[m] < kLaneLow;
        if (K == 2 && big) {
            // one target below bit 6, the other at bit 6 or above
            const uint32_t q0 = g.q[0], q1 = g.q[1];
            const bool low0 = q0 < kLaneLow && q1 >= kLaneLow, low1 = q1 < kLaneLow && q0 >= kLaneLow;
            if (low0 || low1) {
                ValMat<R, 32> mv;
                pack_gate_matrix(std::is_same<R, double>::value ? kC128 : kC64, 2, false, g.re, g.im, mv.v);
                const uint64_t npairs = 1ULL << (n - 1);
                const unsigned pblocks = grid_for(npairs, 256);
                const uint32_t tl = low0 ? q0 : q1, th = low0 ? q1 : q0;
                if (stream)
                    hipLaunchKernelGGL((k_gate_lanes_mixed<R, true>), dim3(pblocks), dim3(256), 0, st, (cv<R>*)psi, npairs,
                                       tl, th, low0 ? 1 : 0, mv);
                else
                    hipLaunchKernelGGL((k_gate_lanes_mixed<R, false>), dim3(pblocks), dim3(256), 0, st, (cv<R>*)psi,
                                       npairs, tl, th, low0 ? 1 : 0, mv);
                return hipGetLastError();
            }
        }
        if (lanes) {
            LaneArgs<K> la;
            la.tmask = 0;
            for (int m = 0; m < K; ++m) {
                la.q[m] = g.q[m];
                la.tmask |= 1u << g.q[m];
            }
            for (int c = 0; c < D; ++c) {
                uint32_t o = 0;
                for (int m = 0; m < K; ++m)
                    if ((c >> (K - 1 - m)) & 1) o |= 1u << g.q[m];
                la.offl[c] = o;
            }
            ValMat<R, 2 * D * D> mv;
            pack_gate_matrix(std::is_same<R, double>::value ? kC128 : kC64, K, false, g.re, g.im, mv.v);
            const uint64_t count = 1ULL << n;
            const unsigned lblocks = grid_for(count, 256);
            if (stream)
                hipLaunchKernelGGL((k_gate_lanes<R, K, true>), dim3(lblocks), dim3(256), 0, st, (cv<R>*)psi, count, la, mv);
            else
                hipLaunchKernelGGL((k_gate_lanes<R, K, false>), dim3(lblocks), dim3(256), 0, st, (cv<R>*)psi, count, la, mv);
            return hipGetLastError();
        }
    }
    if constexpr (K <= 3) {
        ValMat<R, (M3 ? 3 : 2) * D * D> mv;
        pack_gate_matrix(std::is_same<R, double>::value ? kC128 : kC64, K, M3, g.re, g.im, mv.v);
        if (whole)
            hipLaunchKernelGGL((k_gate_direct_val<R, K, M3, true>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, ngroups,
                               a, mv, xm);
        else
            hipLaunchKernelGGL((k_gate_direct_val<R, K, M3, false>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, ngroups,
                               a, mv, xm);
    } else if (whole) {
        hipLaunchKernelGGL((k_gate_direct_ptr<R, K, M3, true>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, ngroups,
                           a, (const R*)g.dmat, xm);
    } else {
        hipLaunchKernelGGL((k_gate_direct_ptr<R, K, M3, false>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, ngroups,
                           a, (const R*)g.dmat, xm);
    }
    return hipGetLastError();
}

template <typename R, int K, bool M3>
hipError_t run_tiled(void* psi, int n, hipStream_t st, const GateDesc& g) {
    TiledArgs<K> a;
    fill_tiled_args<K>(a, n, g.q);
    // one tile per block; a dispatch holds fewer than 2^32 work-items, so
    // more tiles than that take several launches (states above 34 qubits)
    const uint64_t maxb = kMaxDispatchItems / 64;
    for (uint64_t t0 = 0; t0 < a.ntiles; t0 += maxb)
        hipLaunchKernelGGL((k_gate_tiled<R, K, M3>), dim3((unsigned)std::min<uint64_t>(a.ntiles - t0, maxb)), dim3(64), 0,
                           st, (cv<R>*)psi, a, (const R*)g.dmat, t0);
    return hipGetLastError();
}

static int g_mfma_batches = 0;   // 0 = per-precision default
static int g_mfma_lds = -1;      // -1 = per-precision default
static int g_mfma_nt = -1;       // non-temporal loads and stores: 1 on, 0 off, -1 by target set
static int g_mfma_waves = 0;     // waves in the grid (0 = per-precision default)
static int g_mfma_rot = -1;      // super-batch order rotation (-1 = by target set)

// The streaming hint pays only when every load/store instruction covers whole
// 128-byte lines.  One instruction spans the two lowest targets (lane quad) and
// the four lowest non-targets (groups); a line is 3 index bits (c128) or 4
// (c64), so the lines are whole unless three or more targets sit below that.
// With a split line the two halves arrive from two instructions, and the hint
// makes L2 drop the line between them, so it is fetched twice.
inline bool mfma_whole_lines(const MfmaArgs& a, bool f64) {
    const uint32_t line_bits = f64 ? 3u : 4u;
    int low = 0;
    for (int m = 0; m < 5; ++m) low += (a.pos[m] < line_bits);
    return low <= 2;
}

// Launch defaults per target set, from sweeps through the engine at 30
// qubits (dense Haar blocks on the same GPU).  complex128 sets whose
// instructions cover whole lines (the streaming case, 191 of the benchmark's
// 200 blocks): two batches per iteration, 524288 waves, and waves that run
// together taking super-batches two apart (rotation 1).  That gives 5.50 /
// 5.86 TB/s on unrestricted / qubit 0-9 sets, against 5.29 / 5.70 with four
// batches, 131072 waves and the rotation below; the fewer VGPRs per wave let
// more waves share a SIMD.  Sets with three or more targets inside a line
// (not streamed) measured 6.34 ms with four batches, 131072 waves, rotation 2
// when a target lies at qubit 10 or above and 0 otherwise, against 7.09 ms
// with the streaming defaults, so they keep those.  complex64: two batches
// and 524288 waves when a target lies at qubit 10 or above, else four
// batches and 262144 waves, no rotation (any rotation lost, 4.75 vs 5.28 TB/s).  complex64
// lanes moving two adjacent amplitudes per 16-byte access (the even and odd
// groups of two batches, where qubit 0 is not a target) halve the memory
// instructions but measured 5.28 against 5.38 TB/s at every batch count, grid
// and rotation tried, so each lane moves one.  Only large states were swept,
// so the rotation applies from 2^16 super-batches.
struct MfmaDefaults {
    int batches;
    uint64_t waves;
};

inline MfmaDefaults mfma_defaults(const MfmaArgs& a, bool f64) {
    // complex64 at 30 qubits: sets with a target at qubit 10 or above 5.49
    // TB/s with two batches and 524288 waves, 5.42 with 262144 and 5.35 with
    // four batches; sets within qubits 0-9 5.47 with four batches and 262144
    // waves, 5.42 with two
    if (!f64) return a.pos[4] >= 10 ? MfmaDefaults{2, 524288} : MfmaDefaults{4, 262144};
    return mfma_whole_lines(a, f64) ? MfmaDefaults{2, 524288} : MfmaDefaults{4, 131072};
}

inline uint32_t mfma_rotation(const MfmaArgs& a, bool f64, uint64_t nsuper) {
    if (!f64 || nsuper < (1ULL << 16)) return 0u;
    if (mfma_whole_lines(a, f64)) return 1u;
    return a.pos[4] >= 10 ? 2u : 0u;
}

template <typename R, int UB, bool WLDS>
hipError_t launch_mfma(void* psi, const MfmaArgs& args, hipStream_t st, const void* dmat) {
    constexpr bool f64 = std::is_same<R, double>::value;
    MfmaArgs a = args;
    const uint64_t nsuper = a.nbatches / UB;
    uint32_t bits = 0;
    while ((1ULL << bits) < nsuper) ++bits;
    a.bits = ((1ULL << bits) == nsuper) ? bits : 0;  // a rotation needs a power-of-two count
    a.rot = g_mfma_rot >= 0 ? (uint32_t)g_mfma_rot : mfma_rotation(a, f64, nsuper);
    if (a.bits == 0) a.rot = 0;
    // grid: the per-set default (for the streaming c128 case 262144 waves ran
    // 1.4% slower and 1048576 5% slower than 524288)
    const uint64_t target = g_mfma_waves > 0 ? (uint64_t)g_mfma_waves : mfma_defaults(a, f64).waves;
    const uint64_t waves = std::max<uint64_t>(1, std::min<uint64_t>(nsuper, target));
    const uint64_t blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffULL) return hipErrorInvalidValue;
    // small states (up to kXcdStateMax): one super-batch per wave, in the
    // XCD-consistent order of the other gate kernels
    const uint64_t state_bytes = a.nbatches * 16 * 32 * sizeof(cv<R>);  // 2^n amplitudes
    a.xm = (waves == nsuper && a.rot == 0 && a.bits >= 3 && blocks * 4 == nsuper && blocks % kXcds == 0 &&
            state_bytes <= kXcdStateMax)
               ? a.bits
               : 0u;
    // streaming hints only for states larger than the 256 MiB Infinity Cache:
    // a smaller state stays there between passes when cached (24 random
    // 5-qubit blocks, complex128: 22 qubits 37.7 -> 32.7 us per pass, 24
    // qubits 122 -> 101 us; 25 qubits 199 us with the hints, 211 without)
    const bool nt = g_mfma_nt >= 0 ? g_mfma_nt != 0 : (mfma_whole_lines(a, f64) && state_bytes > kCacheStateMax);
    if (nt)
        hipLaunchKernelGGL((k_gate_mfma<R, UB, WLDS, true>), dim3((unsigned)blocks), dim3(256), 0, st, (cv<R>*)psi, a,
                           (const R*)dmat);
    else
        hipLaunchKernelGGL((k_gate_mfma<R, UB, WLDS, false>), dim3((unsigned)blocks), dim3(256), 0, st, (cv<R>*)psi,
                           a, (const R*)dmat);
    return hipGetLastError();
}

template <typename R>
hipError_t run_mfma(void* psi, int n, hipStream_t st, const GateDesc& g) {
    MfmaArgs a;
    uint32_t s_in[32], s_out[32];
    fill_mfma_args(a, n, g.q, std::is_same<R, double>::value, s_in, s_out);
    if (a.nbatches == 0 || !g.dmat) return hipErrorInvalidValue;
    constexpr bool f64 = std::is_same<R, double>::value;
    // batches in flight per wave: the per-set default, fragments in VGPRs
    int ub = g_mfma_batches ? g_mfma_batches : mfma_defaults(a, f64).batches;
    const bool lds = g_mfma_lds >= 0 ? g_mfma_lds != 0 : false;
    while (ub > 1 && a.nbatches % ub) ub >>= 1;
    // small and mid-size states: more waves before more batches per wave,
    // keeping at least 32768 waves in the grid (measured on fused 24-qubit
    // circuits: one batch per wave 12.6 ms, four 14.1 ms for the same blocks)
    if (!g_mfma_batches)
        while (ub > 1 && a.nbatches / ub < 32768) ub >>= 1;
    if (lds) {
        if (ub >= 4) return launch_mfma<R, 4, true>(psi, a, st, g.dmat);
        if (ub == 2) return launch_mfma<R, 2, true>(psi, a, st, g.dmat);
        return launch_mfma<R, 1, true>(psi, a, st, g.dmat);
    }
    if (ub >= 4) return launch_mfma<R, 4, false>(psi, a, st, g.dmat);
    if (ub == 2) return launch_mfma<R, 2, false>(psi, a, st, g.dmat);
    return launch_mfma<R, 1, false>(psi, a, st, g.dmat);
}

template <typename R>
hipError_t dispatch(void* psi, int n, hipStream_t st, const GateDesc& g, int path) {
    const bool m3 = use_three_mult(g.K);
    if (path == kPathMfma && g.K == 5) return run_mfma<R>(psi, n, st, g);
    if (path == kPathTiled) {
        switch (g.K) {
            case 3: return m3 ? run_tiled<R, 3, true>(psi, n, st, g) : run_tiled<R, 3, false>(psi, n, st, g);
            case 4: return m3 ? run_tiled<R, 4, true>(psi, n, st, g) : run_tiled<R, 4, false>(psi, n, st, g);
            case 5: return m3 ? run_tiled<R, 5, true>(psi, n, st, g) : run_tiled<R, 5, false>(psi, n, st, g);
            default: break;
        }
    }
    switch (g.K) {
        case 1: return run_direct<R, 1, false>(psi, n, st, g);
        case 2: return run_direct<R, 2, false>(psi, n, st, g);
        case 3: return m3 ? run_direct<R, 3, true>(psi, n, st, g) : run_direct<R, 3, false>(psi, n, st, g);
        case 4: return m3 ? run_direct<R, 4, true>(psi, n, st, g) : run_direct<R, 4, false>(psi, n, st, g);
        case 5: return m3 ? run_direct<R, 5, true>(psi, n, st, g) : run_direct<R, 5, false>(psi, n, st, g);
        default: return hipErrorInvalidValue;
    }
}

template <typename R, int K>
hipError_t run_diag(void* psi, int n, hipStream_t st, const uint32_t* q, const double* dre, const double* dim) {
    DirectArgs<K> a;
    for (int m = 0; m < K; ++m) a.pos[m] = q[m];  // kept in list order: local bit K-1-m <-> q[m]
    for (int l = 0; l < (1 << K); ++l) a.off[l] = 0;
    ValMat<double, 2 * (1 << K)> d;
    for (int l = 0; l < (1 << K); ++l) {
        d.v[l] = dre[l];
        d.v[(1 << K) + l] = dim[l];
    }
    const uint64_t count = 1ULL << n;
    // one thread per four amplitudes: every load of a thread issued at once
    const unsigned blocks = grid_for((count + kPassSpan - 1) / kPassSpan, 256);
    const uint32_t xm = kPassSpan == 1 ? xcd_log2(count, blocks, count * sizeof(cv<R>)) : 0u;
    if (beyond_cache(count, sizeof(cv<R>)) && g_direct_nt != 0)
        hipLaunchKernelGGL((k_diag<R, K, true>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, count, a, d, xm);
    else
        hipLaunchKernelGGL((k_diag<R, K, false>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, count, a, d, xm);
    return hipGetLastError();
}

}  // namespace

bool stream_hint_enabled() { return g_direct_nt != 0; }

bool use_three_mult(int K) { return K >= 4; }

// Matrix-core kernel options chosen by the autotuner (or left at the defaults).
// `nontemporal` 0 also keeps the direct kernel off the streaming hint.
void set_mfma_options(int batches_per_iteration, int matrix_in_lds, int nontemporal, int grid_waves,
                      int batch_rotation) {
    g_mfma_batches = (batches_per_iteration == 1 || batches_per_iteration == 2 || batches_per_iteration == 4)
                         ? batches_per_iteration : 0;
    g_mfma_lds = (matrix_in_lds == 0 || matrix_in_lds == 1) ? matrix_in_lds : -1;
    g_mfma_nt = (nontemporal == 0 || nontemporal == 1) ? nontemporal : -1;
    g_direct_nt = nontemporal == 0 ? 0 : 1;  // also the whole-state passes and k_diag
    g_mfma_waves = grid_waves > 0 ? grid_waves : 0;
    g_mfma_rot = (batch_rotation >= 0 && batch_rotation < 32) ? batch_rotation : -1;
}

size_t gate_matrix_bytes(int prec, int K, bool three_mult) {
    const size_t d = 1ULL << K;
    return (three_mult ? 3 : 2) * d * d * (prec == kC128 ? sizeof(double) : sizeof(float));
}

// Row r occupies [w*r, w*(r+1)) with w = 2D (A_r, B_r) or 3D (A_r, B_r, S_r),
// so one matrix row is one contiguous run of scalar loads.
void pack_gate_matrix(int prec, int K, bool three_mult, const double* re, const double* im, void* out) {
    const size_t d = 1ULL << K, w = (three_mult ? 3 : 2) * d;
    for (size_t r = 0; r < d; ++r)
        for (size_t c = 0; c < d; ++c) {
            const double a = re[r * d + c], b = im[r * d + c];
            const double v[3] = {a, b, a + b};
            for (int part = 0; part < (three_mult ? 3 : 2); ++part) {
                const size_t at = r * w + part * d + c;
                if (prec == kC128) static_cast<double*>(out)[at] = v[part];
                else static_cast<float*>(out)[at] = (float)v[part];
            }
        }
}

size_t mfma_fragment_bytes(int prec) { return 48 * 64 * (prec == kC128 ? sizeof(double) : sizeof(float)); }

// Matrix-core fragment order for k_gate_mfma: fragment f = (part * 2 + mt) * 8 + kb
// (part 0 = Re U, 1 = Im U, 2 = Re U + Im U) holds, in lane l, the entry
// U_part[s_out(16 mt + l % 16)][s_in(4 kb + l / 16)].
void pack_mfma_fragments(int prec, int n, const uint32_t* q, const double* re, const double* im, void* out) {
    MfmaArgs a;
    uint32_t s_in[32], s_out[32];
    fill_mfma_args(a, n, q, prec == kC128, s_in, s_out);
    for (int part = 0; part < 3; ++part)
        for (int mt = 0; mt < 2; ++mt)
            for (int kb = 0; kb < 8; ++kb)
                for (int l = 0; l < 64; ++l) {
                    const uint32_t r = s_out[16 * mt + (l & 15)], c = s_in[4 * kb + (l >> 4)];
                    const double x = re[r * 32 + c], y = im[r * 32 + c];
                    const double v = part == 0 ? x : (part == 1 ? y : x + y);
                    const size_t at = (size_t)((part * 2 + mt) * 8 + kb) * 64 + l;
                    if (prec == kC128) static_cast<double*>(out)[at] = v;
                    else static_cast<float*>(out)[at] = (float)v;
                }
}

int choose_path(int n, int prec, const GateDesc& g) {
    if (g.path != kPathAuto) return g.path;
    if (g.K == 5 && n >= 9) return kPathMfma;
    if (g.K <= 2 || n < g.K + 6) return kPathDirect;
    uint32_t lowest = 64;
    for (int m = 0; m < g.K; ++m) lowest = std::min(lowest, g.q[m]);
    const uint32_t coalesce_bits = (prec == kC128) ? 3u : 4u;  // 128 B runs
    return lowest >= coalesce_bits ? kPathDirect : kPathTiled;
}

hipError_t launch_gate(void* psi, int n, int prec, hipStream_t st, const GateDesc& g) {
    if (g.K < 1 || g.K > kMaxK || g.K > n) return hipErrorInvalidValue;
    int path = choose_path(n, prec, g);
    if (path == kPathMfma && (g.K != 5 || n < 9)) path = kPathDirect;
    if (path == kPathTiled && (g.K < 3 || n < g.K + 6)) path = kPathDirect;
    if (g.K >= 4 && !g.dmat) return hipErrorInvalidValue;
    if (path == kPathTiled && !g.dmat) path = kPathDirect;
    return prec == kC128 ? dispatch<double>(psi, n, st, g, path) : dispatch<float>(psi, n, st, g, path);
}

namespace {
template <typename R, int K>
hipError_t run_controlled(void* psi, int n, hipStream_t st, const uint32_t* c, uint32_t t, const double* ure,
                          const double* uim) {
    DirectArgs<K> a;
    uint64_t cbits = 0;
    for (int m = 0; m + 1 < K; ++m) {
        a.pos[m] = c[m];
        cbits |= 1ULL << c[m];
    }
    a.pos[K - 1] = t;
    std::sort(a.pos, a.pos + K);
    for (int l = 0; l < (1 << K); ++l) a.off[l] = 0;
    ValMat<R, 8> m;  // row-major 2x2: real parts, then imaginary parts, per row
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 2; ++k) {
            m.v[r * 4 + k] = (R)ure[r * 2 + k];
            m.v[r * 4 + 2 + k] = (R)uim[r * 2 + k];
        }
    const uint64_t count = 1ULL << n, npairs = count >> K;
    const unsigned blocks = grid_for(npairs, 256);
    const uint32_t xm = xcd_log2(npairs, blocks, count * sizeof(cv<R>));
    // the hint by the bytes touched, 2^-(K-1) of the state (complex128 25
    // qubits, one control: 63 us cached, 88 us hinted; 26 qubits 191 / 178)
    if (beyond_cache(count >> (K - 1), sizeof(cv<R>)) && g_direct_nt != 0)
        hipLaunchKernelGGL((k_gate_controlled<R, K, true>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, npairs, a, cbits,
                           1ULL << t, m, xm);
    else
        hipLaunchKernelGGL((k_gate_controlled<R, K, false>), dim3(blocks), dim3(256), 0, st, (cv<R>*)psi, npairs, a,
                           cbits, 1ULL << t, m, xm);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_controlled(void* psi, int n, int prec, hipStream_t st, const uint32_t* c, int nc, uint32_t t,
                             const double* ure, const double* uim) {
    if (nc < 1 || nc > 2 || nc + 1 > n || (int)t >= n) return hipErrorInvalidValue;
    for (int m = 0; m < nc; ++m)
        if ((int)c[m] >= n || c[m] == t || (m == 1 && c[1] == c[0])) return hipErrorInvalidValue;
    if (nc == 1)
        return prec == kC128 ? run_controlled<double, 2>(psi, n, st, c, t, ure, uim)
                             : run_controlled<float, 2>(psi, n, st, c, t, ure, uim);
    return prec == kC128 ? run_controlled<double, 3>(psi, n, st, c, t, ure, uim)
                         : run_controlled<float, 3>(psi, n, st, c, t, ure, uim);
}

hipError_t launch_diag(void* psi, int n, int prec, hipStream_t st, int K, const uint32_t* q, const double* dre,
                       const double* dim) {
    if (K < 1 || K > kMaxK || K > n) return hipErrorInvalidValue;
    if (prec == kC128) {
        switch (K) {
            case 1: return run_diag<double, 1>(psi, n, st, q, dre, dim);
            case 2: return run_diag<double, 2>(psi, n, st, q, dre, dim);
            case 3: return run_diag<double, 3>(psi, n, st, q, dre, dim);
            case 4: return run_diag<double, 4>(psi, n, st, q, dre, dim);
            default: return run_diag<double, 5>(psi, n, st, q, dre, dim);
        }
    }
    switch (K) {
        case 1: return run_diag<float, 1>(psi, n, st, q, dre, dim);
        case 2: return run_diag<float, 2>(psi, n, st, q, dre, dim);
        case 3: return run_diag<float, 3>(psi, n, st, q, dre, dim);
        case 4: return run_diag<float, 4>(psi, n, st, q, dre, dim);
        default: return run_diag<float, 5>(psi, n, st, q, dre, dim);
    }
}

}  // namespace iqg
