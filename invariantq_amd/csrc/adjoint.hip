// Parametric circuits on the GPU: forward evaluation of <H>, adjoint-method
// gradients and (for comparison) parameter-shift gradients.
//
// No reference counterpart: InvariantQ has no gradient code (its VQE,
// invariantq/benchmarks/vqe.py:44-98, is gradient-free COBYLA over
// StateVectorSimulator.run).  Gate matrices follow invariantq/ir/qir.py:90-136
// and Pauli strings follow core/src/statevector.cpp:258-291.
//
// Adjoint method, for U_k = exp(-i theta_k G_k / 2) with Pauli generator G_k:
//   psi_L = U_L..U_1 |0>,  lambda = H psi_L,
//   for k = L..1:  dE/dtheta_k = Im <lambda | G_k psi_k>,  psi <- U_k^dag psi,
//                  lambda <- U_k^dag lambda.
// U3(theta, phi, lam) is taken as RZ(phi) RY(theta) RZ(lam) (equal up to a
// global phase, which leaves <H> and its gradient unchanged).
//
// The tape is compiled into segments, reordering only commuting gates:
//   1q   a run of one-qubit gates on one qubit (<= 4 sub-steps): forward is one
//        pass with the product matrix; reverse is one fused pass that walks the
//        sub-steps backwards, accumulating each generator inner product in
//        FP64 and un-applying the gate to psi and lambda in registers;
//   diag a set of CZ gates (mutually commuting): one sign pass;
//   2q   any other two-qubit gate.
// Each reverse pass moves 4 * 2^n * sizeof(amplitude) bytes, the algorithmic
// minimum for updating two states.  Per-block partial sums of every segment
// are reduced by one final kernel, so the sweep never synchronises with the host.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"
#include "split.hpp"

namespace iqg {

enum GateKind : int { kRX = 0, kRY = 1, kRZ = 2, kU3 = 3, kCZ = 4, kCX = 5, kMat1 = 6, kMat2 = 7 };
enum Gen : int { kGenNone = 0, kGenX = 1, kGenY = 2, kGenZ = 3 };

constexpr int kSubMax = 4;
constexpr int kEdgeMax = 64;

namespace {

struct Seg1qArgs {
    double u[kSubMax][8];  // U^dagger of each sub-step: re00 im00 re01 im01 re10 im10 re11 im11
    int gen[kSubMax];
    int count;
};

struct EdgeList {
    uint32_t a[kEdgeMax];  // 32-bit: indexed by a wave-uniform value, a scalar load
    uint32_t b[kEdgeMax];
    int ne;
};

struct ObsTable {
    const uint64_t* xmask;  // per group
    const int* start;       // group g owns terms [start[g], start[g+1])
    const uint64_t* zy;     // per term
    const int* ny;
    const double* coeff;
    int ngroups;
};

template <typename R>
__device__ __forceinline__ void apply2(const double* u, double& x0r, double& x0i, double& x1r, double& x1i) {
    const double y0r = u[0] * x0r - u[1] * x0i + u[2] * x1r - u[3] * x1i;
    const double y0i = u[0] * x0i + u[1] * x0r + u[2] * x1i + u[3] * x1r;
    const double y1r = u[4] * x0r - u[5] * x0i + u[6] * x1r - u[7] * x1i;
    const double y1i = u[4] * x0i + u[5] * x0r + u[6] * x1i + u[7] * x1r;
    x0r = (double)(R)y0r;
    x0i = (double)(R)y0i;
    x1r = (double)(R)y1r;
    x1i = (double)(R)y1i;
}

// Reverse pass over a run of one-qubit gates on qubit q.  Amplitudes are
// widened to FP64 in registers; each sub-step's result is rounded back to the
// state precision, so the arithmetic matches un-applying the gates one by one.
template <typename R>
__global__ void __launch_bounds__(256) k_adj_seg1q(cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, uint64_t npairs,
                                                   int q, Seg1qArgs a, double* __restrict__ part) {
    double acc[kSubMax] = {0.0, 0.0, 0.0, 0.0};
    const uint64_t bit = 1ULL << q, lo = bit - 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += stride) {
        const uint64_t i0 = ((p & ~lo) << 1) | (p & lo), i1 = i0 | bit;
        const cv<R> A0 = psi[i0], A1 = psi[i1], B0 = lam[i0], B1 = lam[i1];
        double a0r = A0.x, a0i = A0.y, a1r = A1.x, a1i = A1.y;
        double b0r = B0.x, b0i = B0.y, b1r = B1.x, b1i = B1.y;
#pragma unroll
        for (int s = kSubMax - 1; s >= 0; --s) {
            if (s >= a.count) continue;
            const int g = a.gen[s];
            if (g == kGenX) {
                acc[s] += (b0r * a1i - b0i * a1r) + (b1r * a0i - b1i * a0r);
            } else if (g == kGenY) {
                acc[s] += (-b0r * a1r - b0i * a1i) + (b1r * a0r + b1i * a0i);
            } else if (g == kGenZ) {
                acc[s] += (b0r * a0i - b0i * a0r) - (b1r * a1i - b1i * a1r);
            }
            apply2<R>(a.u[s], a0r, a0i, a1r, a1i);
            apply2<R>(a.u[s], b0r, b0i, b1r, b1i);
        }
        cv<R> o;
        o.x = (R)a0r; o.y = (R)a0i; psi[i0] = o;
        o.x = (R)a1r; o.y = (R)a1i; psi[i1] = o;
        o.x = (R)b0r; o.y = (R)b0i; lam[i0] = o;
        o.x = (R)b1r; o.y = (R)b1i; lam[i1] = o;
    }
    block_sum_store<kSubMax>(acc, kSubMax, part);
}

// Product of CZ signs over an edge list, on psi and (optionally) lambda.
template <typename R>
__global__ void __launch_bounds__(256) k_cz_set(cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, uint64_t count, EdgeList e) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint32_t par = 0;
        for (int k = 0; k < e.ne; ++k) par ^= (uint32_t)((i >> e.a[k]) & (i >> e.b[k]) & 1ULL);
        if (par) {
            cv<R> v = psi[i];
            v.x = -v.x;
            v.y = -v.y;
            psi[i] = v;
            if (lam) {
                cv<R> w = lam[i];
                w.x = -w.x;
                w.y = -w.y;
                lam[i] = w;
            }
        }
    }
}

// lambda = H psi and E = Re <psi|H psi> in one pass:
// (H psi)[k] = sum_t c_t phase_t(k ^ x_t) psi[k ^ x_t], phase(j) = i^{nY} (-1)^{popcount(j & (Y|Z))}.
// ACCUM: the table's part of H psi is added to lambda; the energy partials
// are this part's share either way.  SKIP (energy only): a term whose
// right-hand amplitude psi[k ^ x] has |.|^2 < 1e-18 adds nothing, as in the
// reference's <psi|P|psi> (core/src/statevector.cpp:273-275); the skip makes
// the sum slightly non-real, so both parts are kept (Re, Im per block).
template <typename R, bool WRITE, bool ACCUM = false, bool SKIP = false>
__global__ void __launch_bounds__(256) k_apply_h(const cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, uint64_t count,
                                                 ObsTable h, double* __restrict__ part) {
    static_assert(!(SKIP && WRITE), "the skip applies to the energy only");
    double e[2] = {0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        double sr = 0.0, si = 0.0;
        for (int g = 0; g < h.ngroups; ++g) {
            const uint64_t j = k ^ h.xmask[g];
            const cv<R> v = psi[j];
            const double vr = v.x, vi = v.y;
            if (SKIP && vr * vr + vi * vi < 1e-18) continue;
            double cr = 0.0, ci = 0.0;  // sum of c_t * phase_t(j)
            for (int t = h.start[g]; t < h.start[g + 1]; ++t) {
                const double sg = (__popcll(j & h.zy[t]) & 1) ? -h.coeff[t] : h.coeff[t];
                switch (h.ny[t] & 3) {
                    case 0: cr += sg; break;
                    case 1: ci += sg; break;
                    case 2: cr -= sg; break;
                    default: ci -= sg; break;
                }
            }
            sr += cr * vr - ci * vi;
            si += cr * vi + ci * vr;
        }
        const cv<R> p = psi[k];
        e[0] += (double)p.x * sr + (double)p.y * si;
        if (SKIP) e[1] += (double)p.x * si - (double)p.y * sr;
        if (WRITE) {
            cv<R> o;
            if (ACCUM) {
                o = lam[k];
                o.x = (R)((double)o.x + sr);
                o.y = (R)((double)o.y + si);
            } else {
                o.x = (R)sr;
                o.y = (R)si;
            }
            lam[k] = o;
        }
    }
    block_sum_store<2>(e, SKIP ? 2 : 1, part);
}

// ------------------------------------------------------------ tiled stages
//
// A stage (runs of one-qubit gates on any qubits, then at most one CZ set) is
// applied in a few passes over the state instead of one pass per run.  Pass 0
// tiles qubits [0, b): tile index bits = global bits 0..b-1.  Pass k >= 1
// tiles r run bits (global 0..r-1, so every access is a run of 2^r
// amplitudes) plus the group [g0, g0 + m) of high qubits.  Every tile is
// staged in LDS by one workgroup; the runs of the pass's qubits are applied
// there, one LDS sweep per run, and the CZ set (diagonal, from the global
// index) last.  Runs on different qubits commute, so splitting a stage's runs
// by qubit over the passes keeps the stage's product.
constexpr int kPassSegMax = 32;

struct PassSeg {
    double u[8];  // the run's product (forward) or its inverse (reverse sweep), re/im row-major
    // What the register groups apply in a pass split into diagonals and real
    // rotations (see PassDesc::greal): the run's part of the split form, or u.
    double v[8];
    int tbit;     // tile bit of the run's qubit
    int slot;     // gradient slot (reverse)
};

constexpr int kGroupMax = 16;
// Runs per register group: up to 4 qubits for complex64, 3 for complex128.
// With 2^13-amplitude tiles and 512 threads, a group of 4 is one 16-element
// sub-cube per thread; wider groups would leave threads idle.
constexpr int kGroupRuns = 4;

// Runs of a pass sit in the descriptor array sorted by tile bit and cut into
// groups of up to 4 consecutive tile bits, one run per bit: a group is applied
// in registers, one LDS round trip for all its runs.
constexpr int kShiftMax = 16;

struct PassDesc {
    int r, m, g0, ts_bits;
    uint64_t ntiles;
    int nseg, first, ncz, ngroups;
    // 32-bit entries throughout: a byte cannot be a scalar load, and a vector
    // load of a descriptor byte costs a memory round trip per use
    uint32_t gt0[kGroupMax], gsize[kGroupMax], gfirst[kGroupMax];  // first tile bit, runs, first run (in the pass)
    // Bit g set when an earlier group of the pass shares a qubit with group g.
    // With none set and psi stored, the reverse sweep takes every group's
    // overlaps at the pass's end and then moves lambda alone.
    uint32_t gundo;
    // The pass's CZ set (ncz edges), split by where the two qubits sit.  Both
    // in the tile: for each distance s between their tile bits, qmask holds
    // the lower bit of every such edge.  One in the tile: (tile bit, global
    // bit) in cza/czb[0, ncross).  Neither: (global, global) in
    // cza/czb[ncross, ncross + nfixed).
    int ncross, nfixed, nshift;
    uint32_t cza[kEdgeMax], czb[kEdgeMax];
    uint32_t qshift[kShiftMax];
    uint32_t qmask[kShiftMax];
    // Split form (groups with no qubit in common): every run is
    // g diag(1, p) R diag(1, q), R a real rotation.  The group applied first
    // takes the q of every other run as one phase per sub-cube (table tpre),
    // the group applied last their p and the product of the g (table tpost);
    // each group's runs apply v (bit g of greal: v is the real rotation
    // [[v0, -v4], [v4, v0]]).  Tables hold 128 entries for tile bits 0-6, then
    // 64 for bits 7-12, as offsets into the launch's table array (-1: none).
    uint32_t greal;
    int tpre, tpost;
    // Register group: the last group, when it holds exactly tile bits 9 ..
    // 9 + log2(E) - 1 (E = tile / 512 elements per thread).  Each thread's
    // elements tid + 512 u are one sub-cube over those bits, so the group is
    // applied in registers as the tile is stored, one LDS round trip fewer.
    // Applied last in both sweeps (reverse: groups ng-2 .. 0, then it).  -1: none.
    int greg;
};


__device__ __forceinline__ uint64_t pass_global(const PassDesc& d, uint64_t t, uint32_t j) {
    const uint64_t low = j & ((1u << d.r) - 1);
    const uint64_t mid = ((uint64_t)(j >> d.r) & ((1ULL << d.m) - 1)) << d.g0;
    const int gap = d.g0 - d.r;  // global bits [r, g0) come from the tile number's low part
    const uint64_t tlo = t & ((1ULL << gap) - 1), thi = t >> gap;
    return low | mid | (tlo << d.r) | (thi << (d.g0 + d.m));
}

// LDS position of tile element j (low four index bits XOR the next four):
// sixteen lanes reading a register group's sub-cube elements hit distinct
// 8-byte slots whichever tile bits the group holds.
__device__ __forceinline__ uint32_t lds_pos(uint32_t j) { return j ^ ((j >> 4) & 15u); }

// CZ set on tile t: the phase of element j is (-1)^(c ^ <j & M> ^ sum_s
// <j & (j >> s) & qmask_s>), where c (edges between two fixed qubits) and M
// (tile bits paired with a fixed qubit that is set) depend only on the tile.
// The sign is applied as elements move between HBM and LDS, not in a sweep.
struct CzTile {
    uint32_t M, c;
};

__device__ __forceinline__ CzTile cz_setup(const PassDesc& d, uint64_t t) {
    const uint64_t base = pass_global(d, t, 0);
    uint32_t M = 0, c = 0;
    for (int k = 0; k < d.ncross; ++k) M ^= (uint32_t)((base >> d.czb[k]) & 1u) << d.cza[k];
    for (int k = d.ncross; k < d.ncross + d.nfixed; ++k) c ^= (uint32_t)((base >> d.cza[k]) & (base >> d.czb[k]) & 1u);
    return CzTile{(uint32_t)__builtin_amdgcn_readfirstlane(M), (uint32_t)__builtin_amdgcn_readfirstlane(c)};
}

// The first two distance classes are evaluated without a branch (entries past
// nshift have a zero mask: their term is 0); a CZ ring has at most two in a tile.
__device__ __forceinline__ bool cz_odd(const PassDesc& d, CzTile z, uint32_t j) {
    uint32_t par = z.c ^ __popc(j & z.M);
    par ^= __popc(j & (j >> d.qshift[0]) & d.qmask[0]) ^ __popc(j & (j >> d.qshift[1]) & d.qmask[1]);
    for (int k = 2; k < d.nshift; ++k) par ^= __popc(j & (j >> d.qshift[k]) & d.qmask[k]);
    return (par & 1u) != 0;
}

template <typename R>
__device__ __forceinline__ cv<R> negate_if(cv<R> v, bool neg) {
    v.x = neg ? -v.x : v.x;
    v.y = neg ? -v.y : v.y;
    return v;
}

// tile in: 16 elements per thread in flight, every register assigned; with
// `cz`, each element takes its CZ sign on the way
template <typename R>
__device__ __forceinline__ void tile_in(const PassDesc& d, uint64_t t, const cv<R>* __restrict__ src, cv<R>* dst,
                                        uint32_t TS, bool cz = false, CzTile z = CzTile{0, 0}) {
    const uint32_t T = blockDim.x, tid = threadIdx.x;
    for (uint32_t j0 = tid; j0 < TS; j0 += 16 * T) {
        cv<R> v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = src[pass_global(d, t, min(j0 + u * T, TS - 1))];
        if (cz) {
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = negate_if<R>(v[u], cz_odd(d, z, j0 + u * T));
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (j0 + u * T < TS) dst[lds_pos(j0 + u * T)] = v[u];
    }
}

template <typename R>
__device__ __forceinline__ void tile_out(const PassDesc& d, uint64_t t, const cv<R>* src, cv<R>* __restrict__ dst,
                                         uint32_t TS, bool cz = false, CzTile z = CzTile{0, 0}) {
    const uint32_t T = blockDim.x, tid = threadIdx.x;
    if (cz) {
        for (uint32_t j = tid; j < TS; j += T) dst[pass_global(d, t, j)] = negate_if<R>(src[lds_pos(j)], cz_odd(d, z, j));
    } else {
        for (uint32_t j = tid; j < TS; j += T) dst[pass_global(d, t, j)] = src[lds_pos(j)];
    }
}

template <typename R>
__device__ __forceinline__ void mat2(const R* u, cv<R>& a, cv<R>& c) {
    const R y0r = u[0] * a.x - u[1] * a.y + u[2] * c.x - u[3] * c.y;
    const R y0i = u[0] * a.y + u[1] * a.x + u[2] * c.y + u[3] * c.x;
    const R y1r = u[4] * a.x - u[5] * a.y + u[6] * c.x - u[7] * c.y;
    const R y1i = u[4] * a.y + u[5] * a.x + u[6] * c.y + u[7] * c.x;
    a.x = y0r; a.y = y0i; c.x = y1r; c.y = y1i;
}

// (a, c) <- [[co, -si], [si, co]] (a, c): a real rotation of a complex pair.
template <typename R>
__device__ __forceinline__ void rot2(R co, R si, cv<R>& a, cv<R>& c) {
    const cv<R> a0 = a, c0 = c;
    a.x = fma(co, a0.x, -(si * c0.x));
    a.y = fma(co, a0.y, -(si * c0.y));
    c.x = fma(si, a0.x, co * c0.x);
    c.y = fma(si, a0.y, co * c0.y);
}

// Phase of a sub-cube from a split pass's table: the group's own tile bits are
// zero in base, and every diagonal in the table is 1 on a zero bit, so the
// sub-cube's elements share tab[base & 127] * tab[128 + (base >> 7)].
template <typename R>
__device__ __forceinline__ cv<R> cube_phase(const cv<R>* tab, uint32_t base) {
    const cv<R> a = tab[base & 127u], b = tab[128u + (base >> 7)];
    cv<R> o;
    o.x = a.x * b.x - a.y * b.y;
    o.y = a.x * b.y + a.y * b.x;
    return o;
}

template <typename R>
__device__ __forceinline__ cv<R> cmulv(cv<R> v, cv<R> ph) {
    cv<R> o;
    o.x = v.x * ph.x - v.y * ph.y;
    o.y = v.x * ph.y + v.y * ph.x;
    return o;
}

// Tile out with the pass's register group (runs on tile bits 9 .. 9 + UB - 1,
// UB = log2 E): each thread takes its elements tid + 512 u from LDS, applies
// the group's runs (v; real rotations when REAL) to the sub-cube, then the
// post table's phase (one value: the sub-cube's base is tid) and the CZ signs,
// and stores.
template <typename R, bool REAL>
__device__ __forceinline__ void tile_out_reg(const PassDesc& d, uint64_t t, const cv<R>* src, cv<R>* __restrict__ dst,
                                             const PassSeg* __restrict__ sg, const cv<R>* post, bool cz, CzTile z) {
    constexpr int E = sizeof(R) == 4 ? 16 : 8, UB = sizeof(R) == 4 ? 4 : 3;
    const uint32_t T = blockDim.x, tid = threadIdx.x, pt = lds_pos(tid);
    R u[UB][REAL ? 2 : 8];
#pragma unroll
    for (int i = 0; i < UB; ++i) {
        if (REAL) {
            u[i][0] = (R)sg[i].v[0];
            u[i][1] = (R)sg[i].v[4];
        } else {
#pragma unroll
            for (int e = 0; e < (REAL ? 2 : 8); ++e) u[i][e] = (R)sg[i].v[e];
        }
    }
    cv<R> x[E];
#pragma unroll
    for (int k = 0; k < E; ++k) x[k] = src[pt + (uint32_t)k * T];
#pragma unroll
    for (int i = 0; i < UB; ++i)
#pragma unroll
        for (int k = 0; k < E; ++k)
            if (!(k & (1 << i))) {
                if (REAL) rot2<R>(u[i][0], u[i][1], x[k], x[k | (1 << i)]);
                else mat2<R>(u[i], x[k], x[k | (1 << i)]);
            }
    if (post) {
        const cv<R> ph = cube_phase<R>(post, tid);
#pragma unroll
        for (int k = 0; k < E; ++k) x[k] = cmulv<R>(x[k], ph);
    }
    uint32_t me = tid;
    asm volatile("" : "+v"(me));  // per-element terms recomputed per tile, not held across the tile loop
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const uint32_t j = me + (uint32_t)k * T;
        dst[pass_global(d, t, j)] = cz ? negate_if<R>(x[k], cz_odd(d, z, j)) : x[k];
    }
}


// The G runs of a group (tile bits t0 .. t0+G-1) on sub-cubes of 2^G elements
// held in registers: each run's v, real rotations when REAL; with `pre` the
// sub-cube first takes its phase from that table, with `post` afterwards.
template <typename R, int G, bool REAL>
__device__ __forceinline__ void group_apply(cv<R>* tile, uint32_t TS, int t0, const PassSeg* __restrict__ sg,
                                            const cv<R>* pre, const cv<R>* post) {
    constexpr uint32_t D = 1u << G;
    const uint32_t T = blockDim.x, tid = threadIdx.x, cubes = TS >> G, lo = (1u << t0) - 1;
    R u[G][REAL ? 2 : 8];
#pragma unroll
    for (int i = 0; i < G; ++i) {
        if (REAL) {
            u[i][0] = (R)sg[i].v[0];
            u[i][1] = (R)sg[i].v[4];
        } else {
#pragma unroll
            for (int e = 0; e < (REAL ? 2 : 8); ++e) u[i][e] = (R)sg[i].v[e];
        }
    }
    for (uint32_t s = tid; s < cubes; s += T) {
        const uint32_t base = ((s & ~lo) << G) | (s & lo);
        const uint32_t pb = lds_pos(base);  // lds_pos is linear over disjoint bits
        cv<R> x[D];
#pragma unroll
        for (uint32_t m = 0; m < D; ++m) x[m] = tile[(pb ^ lds_pos(m << t0))];
        if (pre) {
            const cv<R> ph = cube_phase<R>(pre, base);
#pragma unroll
            for (uint32_t m = 0; m < D; ++m) x[m] = cmulv<R>(x[m], ph);
        }
#pragma unroll
        for (int i = 0; i < G; ++i)
#pragma unroll
            for (uint32_t m = 0; m < D; ++m)
                if (!(m & (1u << i))) {
                    if (REAL) rot2<R>(u[i][0], u[i][1], x[m], x[m | (1u << i)]);
                    else mat2<R>(u[i], x[m], x[m | (1u << i)]);
                }
        if (post) {
            const cv<R> ph = cube_phase<R>(post, base);
#pragma unroll
            for (uint32_t m = 0; m < D; ++m) x[m] = cmulv<R>(x[m], ph);
        }
#pragma unroll
        for (uint32_t m = 0; m < D; ++m) tile[(pb ^ lds_pos(m << t0))] = x[m];
    }
}

// Reverse: the runs of a group sit on distinct qubits and commute, so every
// run's overlaps T_s = Im<lambda| s psi> (s = X, Y, Z on the run's qubit) are
// taken from the same register contents, before any inverse; then each run's
// combined inverse U^dagger is applied to psi and lambda.  A sub-step k of
// the run has gradient Im<lambda| G'_k psi> with G'_k = V_k G_k V_k^dagger
// (V_k the run's later sub-steps), a real combination of the three T_s formed
// on the host.  Products are summed over a sub-cube in the state's precision,
// then into FP64 per-thread sums.
template <typename R, int G>
__device__ __forceinline__ void sub_overlaps(const cv<R> (&x)[1 << G], const cv<R> (&y)[1 << G], double (&acc)[kGroupRuns][3]) {
    constexpr uint32_t D = 1u << G;
    // T_Z of every run from one product per element, w_m = Im(conj(y_m) x_m):
    // T_Z(i) = sum_m (-1)^(bit i of m) w_m = W - 2 (sum of w_m with bit i set)
    R w[D], wall;
#pragma unroll
    for (uint32_t m = 0; m < D; ++m) {
        w[m] = fma(y[m].x, x[m].y, -(y[m].y * x[m].x));
        wall = m ? wall + w[m] : w[m];
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const uint32_t bit = 1u << i;
        // one fused multiply-add per product term, chained through the sum
        // (no separate product, difference and accumulate per pair)
        R tx = 0, ty = 0, w1 = 0;
        bool first = true;
#pragma unroll
        for (uint32_t m = 0; m < D; ++m) {
            if (m & bit) continue;
            const cv<R>&a0 = x[m], &a1 = x[m | bit], &b0 = y[m], &b1 = y[m | bit];
            tx = first ? b0.x * a1.y : fma(b0.x, a1.y, tx);
            ty = first ? b1.x * a0.x : fma(b1.x, a0.x, ty);
            w1 = first ? w[m | bit] : w1 + w[m | bit];
            first = false;
            tx = fma(-b0.y, a1.x, tx);
            tx = fma(b1.x, a0.y, tx);
            tx = fma(-b1.y, a0.x, tx);
            ty = fma(b1.y, a0.y, ty);
            ty = fma(-b0.x, a1.x, ty);
            ty = fma(-b0.y, a1.y, ty);
        }
        acc[i][0] += (double)tx;
        acc[i][1] += (double)ty;
        acc[i][2] += (double)(wall - 2 * w1);
    }
}

template <typename R, int G>
__device__ __forceinline__ void load_runs(const PassSeg* __restrict__ sg, R (&u)[G][8]) {
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
        for (int e = 0; e < 8; ++e) u[i][e] = (R)sg[i].u[e];
}

template <typename R, int G>
__device__ __forceinline__ void rev_group(cv<R>* tp, cv<R>* tl, uint32_t TS, int t0, const PassSeg* __restrict__ sg,
                                          double (&acc)[kGroupRuns][3]) {
    constexpr uint32_t D = 1u << G;
    const uint32_t T = blockDim.x, tid = threadIdx.x, cubes = TS >> G, lo = (1u << t0) - 1;
    R u[G][8];
    load_runs<R, G>(sg, u);
    for (uint32_t s = tid; s < cubes; s += T) {
        const uint32_t base = ((s & ~lo) << G) | (s & lo);
        const uint32_t pb = lds_pos(base);  // lds_pos is linear over disjoint bits
        cv<R> x[D], y[D];
#pragma unroll
        for (uint32_t m = 0; m < D; ++m) {
            x[m] = tp[(pb ^ lds_pos(m << t0))];
            y[m] = tl[(pb ^ lds_pos(m << t0))];
        }
        sub_overlaps<R, G>(x, y, acc);
#pragma unroll
        for (int i = 0; i < G; ++i)
#pragma unroll
            for (uint32_t m = 0; m < D; ++m) {
                if (m & (1u << i)) continue;
                mat2<R>(u[i], x[m], x[m | (1u << i)]);
                mat2<R>(u[i], y[m], y[m | (1u << i)]);
            }
#pragma unroll
        for (uint32_t m = 0; m < D; ++m) {
            tp[(pb ^ lds_pos(m << t0))] = x[m];
            tl[(pb ^ lds_pos(m << t0))] = y[m];
        }
    }
}

// Overlaps of a group's runs only: psi and lambda are read, not changed.
template <typename R, int G>
__device__ __forceinline__ void rev_overlaps(const cv<R>* tp, const cv<R>* tl, uint32_t TS, int t0,
                                             double (&acc)[kGroupRuns][3]) {
    constexpr uint32_t D = 1u << G;
    const uint32_t T = blockDim.x, tid = threadIdx.x, cubes = TS >> G, lo = (1u << t0) - 1;
    for (uint32_t s = tid; s < cubes; s += T) {
        const uint32_t base = ((s & ~lo) << G) | (s & lo);
        const uint32_t pb = lds_pos(base);  // lds_pos is linear over disjoint bits
        cv<R> x[D], y[D];
#pragma unroll
        for (uint32_t m = 0; m < D; ++m) {
            x[m] = tp[(pb ^ lds_pos(m << t0))];
            y[m] = tl[(pb ^ lds_pos(m << t0))];
        }
        sub_overlaps<R, G>(x, y, acc);
    }
}

// Group width dispatch (register groups hold up to 4 complex64 or 3
// complex128 qubits; the complex128 instantiation of 4 is never run).
#define IQ_GROUP_SWITCH(gs, CALL)                                          \
    switch (gs) {                                                          \
        case 1: CALL(1); break;                                            \
        case 2: CALL(2); break;                                            \
        case 3: CALL(3); break;                                            \
        default: if (sizeof(R) == 4) CALL(sizeof(R) == 4 ? 4 : 3); break;  \
    }

// A pass's group g of a split-form sweep: its runs' v (real rotations when
// greal says so), the pre table when g is applied first, the post table when
// last (with one group, neither: the pass is not split).
#define IQ_APPLY_GROUP(tile, g, first_g, last_g)                                                           \
    do {                                                                                                   \
        const cv<R>* pre_ = ((g) == (first_g) && d.tpre >= 0) ? ltab : nullptr;                            \
        const cv<R>* post_ = ((g) == (last_g) && d.tpost >= 0) ? ltab + kPhaseTab : nullptr;                \
        const PassSeg* sg_ = ps + d.gfirst[g];                                                             \
        if ((d.greal >> (g)) & 1u) {                                                                       \
            IQ_GROUP_SWITCH(d.gsize[g], IQ_GA_REAL)                                                        \
        } else {                                                                                           \
            IQ_GROUP_SWITCH(d.gsize[g], IQ_GA_CPLX)                                                        \
        }                                                                                                  \
    } while (0)
#define IQ_GA_REAL(G) group_apply<R, G, true>(tile_, TS, d.gt0[g_], sg_, pre_, post_)
#define IQ_GA_CPLX(G) group_apply<R, G, false>(tile_, TS, d.gt0[g_], sg_, pre_, post_)

// A split pass's phase tables (pre, then post) from the launch's table array
// into LDS, in the state's precision.
template <typename R>
__device__ __forceinline__ void load_phase_tables(const PassDesc& d, const double2* __restrict__ tabs, cv<R>* ltab) {
    if (d.tpre < 0 && d.tpost < 0) return;
    for (int i = threadIdx.x; i < 2 * kPhaseTab; i += blockDim.x) {
        const int off = i < kPhaseTab ? d.tpre : d.tpost;
        if (off < 0) continue;
        const double2 v = tabs[off + i % kPhaseTab];
        cv<R> o;
        o.x = (R)v.x;
        o.y = (R)v.y;
        ltab[i] = o;
    }
}

template <typename R>
__global__ void __launch_bounds__(512) k_tile_fwd(const cv<R>* __restrict__ src, cv<R>* __restrict__ dst, PassDesc d,
                                                  const PassSeg* __restrict__ segs, const double2* __restrict__ tabs) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t TS = 1u << d.ts_bits;
    cv<R>* tile = reinterpret_cast<cv<R>*>(smem);
    cv<R>* ltab = tile + TS;  // [2][kPhaseTab]
    const PassSeg* ps = segs + d.first;
    load_phase_tables<R>(d, tabs, ltab);  // read after the first tile's barrier
    for (uint64_t t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        tile_in<R>(d, t, src, tile, TS);
        __syncthreads();
        const int nlds = d.greg >= 0 ? d.ngroups - 1 : d.ngroups;  // the register group is the last
        for (int gi = 0; gi < nlds; ++gi) {
            // a wave-uniform index: the descriptor reads below become scalar loads
            const int g_ = __builtin_amdgcn_readfirstlane(gi);
            cv<R>* tile_ = tile;
            IQ_APPLY_GROUP(tile_, g_, 0, d.ngroups - 1);
            __syncthreads();
        }
        // the stage's CZ set comes last: its signs go on with the store
        const CzTile z = d.ncz ? cz_setup(d, t) : CzTile{0, 0};
        if (d.greg >= 0) {
            const int g = d.greg;
            const cv<R>* post = d.tpost >= 0 ? ltab + kPhaseTab : nullptr;
            if ((d.greal >> g) & 1u) tile_out_reg<R, true>(d, t, tile, dst, ps + d.gfirst[g], post, d.ncz != 0, z);
            else tile_out_reg<R, false>(d, t, tile, dst, ps + d.gfirst[g], post, d.ncz != 0, z);
        } else {
            tile_out<R>(d, t, tile, dst, TS, d.ncz != 0, z);
        }
        __syncthreads();
    }
}


// Wave sums of a group's overlaps (up to 4 runs x 3 values): a
// reduce-scatter butterfly over the 12 values.  At offset 32 each lane keeps
// one half (6 values) and adds the partner lane's copy of it, at offset 16
// one half of those (3); each of the 3 is then summed over offsets 8..1 and
// added into the wave's slot by the lane with bits 3..0 clear.  6 + 3 + 12
// shuffles per group instead of 6 per value.
template <typename R>
__device__ __forceinline__ void flush_overlaps(const double (&acc)[kGroupRuns][3], int gs, double* wred, int wave, int lane,
                                               int first) {
    double v[6];
    const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double a = acc[i / 3][i % 3], b = acc[2 + i / 3][i % 3];  // values i and 6 + i
        v[i] = (h5 ? b : a) + __shfl_xor(h5 ? a : b, 32, 64);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double a = v[i], b = v[3 + i];
        double x = (h4 ? b : a) + __shfl_xor(h4 ? a : b, 16, 64);
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        const int k = (h5 ? 6 : 0) + (h4 ? 3 : 0) + i;  // value k = 3 run + component
        if ((lane & 15) == 0 && k < 3 * gs) wred[(wave * kPassSegMax + first + k / 3) * kSubMax + k % 3] += x;
    }
}

// Reverse of a forward pass on psi and lambda: the CZ set first, then the
// groups last to first.  STORED: psi is the pass's stored output, only read
// (when a qubit repeats in the pass, its tile copy is taken back alongside
// lambda); otherwise psi is taken back through the pass in place.  Each run's overlaps (T_X, T_Y, T_Z in entries 0-2 of
// its slot, entry 3 zero) go per wave into LDS, per block (fixed order over
// waves) into the partial layout k_finalize_slots reduces; the grid is
// exactly `nblk` blocks, each writing every run's partials, zeros included.
template <typename R, bool STORED>
__global__ void __launch_bounds__(512) k_tile_rev(cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, PassDesc d,
                                                  const PassSeg* __restrict__ segs, const double2* __restrict__ tabs,
                                                  double* __restrict__ gpart, int nblk) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t T = blockDim.x, tid = threadIdx.x, TS = 1u << d.ts_bits;
    const int lane = tid & 63, wave = tid >> 6, nw = T >> 6;
    cv<R>* tp = reinterpret_cast<cv<R>*>(smem);
    cv<R>* tl = tp + TS;
    double* wred = reinterpret_cast<double*>(tl + TS);  // [nw][kPassSegMax][kSubMax]
    cv<R>* ltab = reinterpret_cast<cv<R>*>(wred + nw * kPassSegMax * kSubMax);  // [2][kPhaseTab]
    for (int i = tid; i < nw * kPassSegMax * kSubMax; i += T) wred[i] = 0.0;
    if (STORED && d.gundo == 0) load_phase_tables<R>(d, tabs, ltab);
    __syncthreads();
    const PassSeg* ps = segs + d.first;
    for (uint64_t t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        // the CZ set is undone first: its signs go on with the loads
        const CzTile z = d.ncz ? cz_setup(d, t) : CzTile{0, 0};
        tile_in<R>(d, t, psi, tp, TS, d.ncz != 0, z);
        tile_in<R>(d, t, lam, tl, TS, d.ncz != 0, z);
        __syncthreads();
        if (STORED && d.gundo == 0) {
            // no qubit repeats in the pass: every group's overlaps at the
            // pass's end (psi as stored), then lambda back through the groups
            for (int gi = d.ngroups - 1; gi >= 0; --gi) {
                // a wave-uniform index: the descriptor reads below become scalar loads
                const int g = __builtin_amdgcn_readfirstlane(gi);
                const int t0 = d.gt0[g], gs = d.gsize[g];
                double acc[kGroupRuns][3] = {};
#define IQ_CALL(G) rev_overlaps<R, G>(tp, tl, TS, t0, acc)
                IQ_GROUP_SWITCH(gs, IQ_CALL)
#undef IQ_CALL
                flush_overlaps<R>(acc, gs, wred, wave, lane, d.gfirst[g]);
            }
            __syncthreads();
            // lambda back: groups ng-1 .. 0, or with a register group ng-2 .. 0
            // and then it (the order the split form's tables assume)
            const int nlds = d.greg >= 0 ? d.ngroups - 1 : d.ngroups;
            const int fg = nlds - 1 >= 0 ? nlds - 1 : d.greg, lg = d.greg >= 0 ? d.greg : 0;
            for (int gi = nlds - 1; gi >= 0; --gi) {
                const int g_ = __builtin_amdgcn_readfirstlane(gi);
                cv<R>* tile_ = tl;
                IQ_APPLY_GROUP(tile_, g_, fg, lg);
                __syncthreads();
            }
            if (d.greg >= 0) {
                const int g_ = d.greg;
                cv<R>* tile_ = tl;
                IQ_APPLY_GROUP(tile_, g_, fg, lg);
                __syncthreads();
            }
        } else {
            // psi and lambda back together, group by group (a repeated qubit
            // needs the cut between its runs)
            for (int gi = d.ngroups - 1; gi >= 0; --gi) {
                const int g = __builtin_amdgcn_readfirstlane(gi);
                const int t0 = d.gt0[g], gs = d.gsize[g];
                const PassSeg* sg = ps + d.gfirst[g];
                double acc[kGroupRuns][3] = {};
#define IQ_CALL(G) rev_group<R, G>(tp, tl, TS, t0, sg, acc)
                IQ_GROUP_SWITCH(gs, IQ_CALL)
#undef IQ_CALL
                flush_overlaps<R>(acc, gs, wred, wave, lane, d.gfirst[g]);
                __syncthreads();
            }
        }
        if (!STORED) tile_out<R>(d, t, tp, psi, TS);
        tile_out<R>(d, t, tl, lam, TS);
        __syncthreads();
    }
    for (int i = tid; i < d.nseg * kSubMax; i += T) {
        const int k = i / kSubMax, st = i % kSubMax;
        double v = 0.0;
        for (int w = 0; w < nw; ++w) v += wred[(w * kPassSegMax + k) * kSubMax + st];
        gpart[((uint64_t)ps[k].slot * nblk + blockIdx.x) * kSubMax + st] = v;
    }
}

// Wave sums of one run's three overlap sums (FP32 per lane): a reduce-scatter
// over (x, y, z, 0) at offsets 32 and 16 leaves one value per lane, offsets
// 8..1 finish it, and lanes 0, 16, 32, 48 add their component to the wave's
// FP64 slot (component 3 is the padding, zero).
template <typename R>
__device__ __forceinline__ void flush_run(R tx, R ty, R tz, double* slot, int lane) {
    const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0;
    const R a0 = (h5 ? tz : tx) + __shfl_xor(h5 ? tx : tz, 32, 64);
    const R a1 = (h5 ? R(0) : ty) + __shfl_xor(h5 ? ty : R(0), 32, 64);
    R v = (h4 ? a1 : a0) + __shfl_xor(h4 ? a0 : a1, 16, 64);
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int c = (h5 ? 2 : 0) + (h4 ? 1 : 0);  // 0 x, 1 y, 2 z, 3 padding
    if ((lane & 15) == 0 && c < 3) slot[c] += (double)v;
}

// Reverse of a forward pass whose output psi was stored and whose groups sit
// on distinct qubits (no group undoes an earlier one): lambda alone is staged
// in LDS, half the footprint of staging psi too, so two workgroups share a
// compute unit and one's loads overlap the other's arithmetic.
//   Overlaps: each thread reads its E = TS / 512 elements j = tid + 512 u of
//   psi (the stored pass output) and lambda once, in one coalesced round,
//   with the pass's CZ signs; lambda goes to LDS.  For every run, on tile bit
//   b, the thread's terms pair its own psi_j with lambda_{j ^ 2^b} from LDS:
//     T_X = sum_j Im(conj(lambda_{j^b}) psi_j),
//     T_Y = sum_j (-1)^{j_b} Re(conj(lambda_{j^b}) psi_j),
//     T_Z = sum_j (-1)^{j_b} Im(conj(lambda_j) psi_j),
//   (the runs commute, so all are taken at the pass's end).  For b < 9 the
//   sign is the thread's own; for b >= 9 it is u's, the same for every lane.
//   Then lambda goes back through the groups.  Partials as k_tile_rev.
template <typename R>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_tile_rev_lam(const cv<R>* __restrict__ psi, cv<R>* __restrict__ lam,
                                                      PassDesc d, const PassSeg* __restrict__ segs,
                                                      const double2* __restrict__ tabs,
                                                      double* __restrict__ gpart, int nblk) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int E = sizeof(R) == 4 ? 16 : 8;  // elements per thread: 2^13 (2^12) tiles over 512 threads
    const uint32_t T = blockDim.x, tid = threadIdx.x, TS = 1u << d.ts_bits;
    const int lane = tid & 63, wave = tid >> 6, nw = T >> 6;
    cv<R>* tl = reinterpret_cast<cv<R>*>(smem);
    double* wred = reinterpret_cast<double*>(tl + TS);  // [nw][kPassSegMax][kSubMax]
    cv<R>* ltab = reinterpret_cast<cv<R>*>(wred + nw * kPassSegMax * kSubMax);  // [2][kPhaseTab]
    for (int i = tid; i < nw * kPassSegMax * kSubMax; i += T) wred[i] = 0.0;
    load_phase_tables<R>(d, tabs, ltab);
    __syncthreads();
    const PassSeg* ps = segs + d.first;
    const uint32_t pt = lds_pos(tid);  // element tid + 512 u sits at pt + 512 u (lds_pos moves bits 4-7 only)
    for (uint64_t t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        cv<R> x[E];  // psi_j, j = tid + 512 u
        // T_Z sums: wall = sum_u w_u, wz[i] = sum_u (-1)^{u_i} w_u (tile bit 9 + i),
        // w_u = Im(conj(lambda_j) psi_j)
        constexpr int UB = E == 16 ? 4 : 3;
        R wall = 0, wz[UB];
#pragma unroll
        for (int i = 0; i < UB; ++i) wz[i] = 0;
        {
            const CzTile z = d.ncz ? cz_setup(d, t) : CzTile{0, 0};
            // the element indices from an opaque copy of tid: per-element
            // terms that do not depend on the tile are recomputed each tile,
            // not kept live (spilled) across the tile loop
            uint32_t me = tid;
            asm volatile("" : "+v"(me));
            // two rounds of E / 2 elements per array in flight (registers)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                cv<R> y[E / 2];
#pragma unroll
                for (int v = 0; v < E / 2; ++v) {
                    const int u = h * (E / 2) + v;
                    const uint64_t g = pass_global(d, t, me + (uint32_t)u * T);
                    x[u] = psi[g];
                    y[v] = lam[g];
                }
#pragma unroll
                for (int v = 0; v < E / 2; ++v) {
                    const int u = h * (E / 2) + v;
                    if (d.ncz) {
                        const bool odd = cz_odd(d, z, me + (uint32_t)u * T);
                        x[u] = negate_if<R>(x[u], odd);
                        y[v] = negate_if<R>(y[v], odd);
                    }
                    const R w = fma(y[v].x, x[u].y, -(y[v].y * x[u].x));
                    wall += w;
#pragma unroll
                    for (int i = 0; i < UB; ++i) wz[i] += ((u >> i) & 1) ? -w : w;
                    tl[pt + (uint32_t)u * T] = y[v];
                }
            }
        }
        __syncthreads();
        for (int k = 0; k < d.nseg; ++k) {
            // a wave-uniform index: the run's tile bit is a scalar load
            const int kk = __builtin_amdgcn_readfirstlane(k);
            const int b = ps[kk].tbit;
            R tx, ty, tz;
            if (b < 9) {  // partner in another thread, same u; sign from the thread's own bit
                const uint32_t q = pt ^ lds_pos(1u << b);
#pragma unroll
                for (int u = 0; u < E; ++u) {
                    const cv<R> l = tl[q + (uint32_t)u * T];
                    tx = u ? fma(l.x, x[u].y, tx) : l.x * x[u].y;
                    tx = fma(-l.y, x[u].x, tx);
                    ty = u ? fma(l.x, x[u].x, ty) : l.x * x[u].x;
                    ty = fma(l.y, x[u].y, ty);
                }
                const bool up = (tid >> b) & 1u;
                ty = up ? -ty : ty;
                tz = up ? -wall : wall;
            } else {  // partner u ^ 2^(b-9) of this thread; the sign is u's bit, the same on every lane
                const uint32_t ub = 1u << (b - 9);
                tx = ty = 0;
#pragma unroll
                for (int u = 0; u < E; ++u) {
                    const cv<R> l = tl[pt + ((uint32_t)u ^ ub) * T];
                    const R sg = ((uint32_t)u & ub) ? R(-1) : R(1);
                    tx = fma(l.x, x[u].y, tx);
                    tx = fma(-l.y, x[u].x, tx);
                    ty = fma(sg, fma(l.x, x[u].x, l.y * x[u].y), ty);
                }
                tz = wz[0];
#pragma unroll
                for (int i = 1; i < UB; ++i) tz = (b - 9 == i) ? wz[i] : tz;
            }
            flush_run<R>(tx, ty, tz, wred + (wave * kPassSegMax + kk) * kSubMax, lane);
        }
        __syncthreads();
        // lambda back: groups ng-1 .. 0, or with a register group ng-2 .. 0 in
        // LDS and then it in registers as the tile is stored
        const int nlds = d.greg >= 0 ? d.ngroups - 1 : d.ngroups;
        const int fg = nlds - 1 >= 0 ? nlds - 1 : d.greg, lg = d.greg >= 0 ? d.greg : 0;
        for (int gi = nlds - 1; gi >= 0; --gi) {
            const int g_ = __builtin_amdgcn_readfirstlane(gi);
            cv<R>* tile_ = tl;
            IQ_APPLY_GROUP(tile_, g_, fg, lg);
            __syncthreads();
        }
        if (d.greg >= 0) {
            const int g = d.greg;
            const cv<R>* post = (g == lg && d.tpost >= 0) ? ltab + kPhaseTab : nullptr;
            if ((d.greal >> g) & 1u) tile_out_reg<R, true>(d, t, tl, lam, ps + d.gfirst[g], post, false, CzTile{0, 0});
            else tile_out_reg<R, false>(d, t, tl, lam, ps + d.gfirst[g], post, false, CzTile{0, 0});
        } else {
            tile_out<R>(d, t, tl, lam, TS);
        }
        __syncthreads();
    }
    for (int i = tid; i < d.nseg * kSubMax; i += T) {
        const int k = i / kSubMax, st = i % kSubMax;
        double v = 0.0;
        for (int w = 0; w < nw; ++w) v += wred[(w * kPassSegMax + k) * kSubMax + st];
        gpart[((uint64_t)ps[k].slot * nblk + blockIdx.x) * kSubMax + st] = v;
    }
}

// ------------------------------------------------------ tiled observable
// The terms of H are grouped by flip mask x; a group whose x lies on the tile
// bits of some pass is applied there: psi's tile is staged in LDS, the
// partner of element j is element j ^ x_tile, and a term's sign bits split
// into tile bits (zl) and bits fixed by the tile index (zfix, one parity per
// tile and term).  Every launch adds its groups' part of H psi to lambda (the
// first one writes it) and leaves its part of Re <psi|H psi> in `part`.
// A group's terms come as two runs: [first, first + nre) add to the real
// part of the phase sum (i^nY = +-1), the rest to the imaginary part (i^nY =
// +-i); the sign of i^nY is folded into the coefficient.  Each run starts
// with its terms that have no sign bits on the tile (nre0 and nim0 of them),
// whose sum is one value per tile.
struct ObsGroupT {
    uint32_t xl;  // flip mask on tile bits
    int first, count, nre;  // terms, relative to the pass's first term
    int nre0, nim0;
    // After the real run's nre0 tile-constant terms come nch terms that the
    // energy kernel takes as ccount chain classes (table index cfirst); the
    // lambda kernel takes them term by term.
    int nch, cfirst, ccount;
};

// A chain class: real-part terms with the same coefficient whose sign bits on
// the tile are one bit a (shift 0) or two bits a, a + shift, and none fixed
// by the tile number.  mask holds every such a (distinct), so for element j
// the class adds coeff * (count - 2 popcount(f(j) & mask)), f(j) = j for
// shift 0 and j ^ (j >> shift) otherwise: a uniform ZZ chain in one popcount.
struct ObsChainT {
    uint32_t mask;
    int shift, count, pad;
    double coeff;
};

struct ObsTermT {
    uint64_t zfix;  // sign bits outside the tile (global positions)
    uint32_t zl;    // sign bits on tile bits
    double coeff;   // times the sign of i^nY
};

constexpr int kObsTermsMax = 2048;  // terms per launch (their per-tile coefficients sit in LDS)

struct ObsPass {
    PassDesc d;  // geometry only
    int gfirst, gcount, tfirst, tcount;
};

template <typename R, bool WRITE, bool FIRST>
__global__ void __launch_bounds__(512) k_obs_tile(const cv<R>* __restrict__ psi, cv<R>* __restrict__ lam, ObsPass op,
                                                  const ObsGroupT* __restrict__ groups,
                                                  const ObsTermT* __restrict__ terms, double* __restrict__ part) {
    // Each thread holds E tile elements (j = chunk + e T + tid) in registers, so
    // a term's data (read from LDS) serves E elements: the term loop is
    // arithmetic on independent elements instead of a chain of table loads.
    constexpr int E = 8;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double wsum[16];
    const PassDesc& d = op.d;
    const uint32_t T = blockDim.x, tid = threadIdx.x, TS = 1u << d.ts_bits;
    cv<R>* tile = reinterpret_cast<cv<R>*>(smem);
    double* csg = reinterpret_cast<double*>(tile + TS);  // each term's coefficient with its tile sign
    uint32_t* zls = reinterpret_cast<uint32_t*>(csg + op.tcount);  // each term's sign bits on tile bits
    const ObsGroupT* gp = groups + op.gfirst;
    const ObsTermT* tp = terms + op.tfirst;
    for (int k = tid; k < op.tcount; k += T) zls[k] = tp[k].zl;
    double e = 0.0;
    for (uint64_t t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        tile_in<R>(d, t, psi, tile, TS);
        const uint64_t base = pass_global(d, t, 0);
        for (int k = tid; k < op.tcount; k += T) csg[k] = (__popcll(base & tp[k].zfix) & 1) ? -tp[k].coeff : tp[k].coeff;
        __syncthreads();
        for (uint32_t c0 = 0; c0 < TS; c0 += E * T) {
            uint32_t jj[E];
            double sr[E], si[E];
#pragma unroll
            for (int q = 0; q < E; ++q) {
                jj[q] = c0 + q * T + tid;
                sr[q] = 0.0;
                si[q] = 0.0;
            }
            for (int gi = 0; gi < op.gcount; ++gi) {
                // a wave-uniform index: the group table reads below become scalar loads
                const int g = __builtin_amdgcn_readfirstlane(gi);
                const uint32_t xl = gp[g].xl;
                double vr[E], vi[E], cr[E], ci[E];  // partner amplitude; sum of c_t * phase_t(partner)
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    vr[q] = vi[q] = 0.0;
                    if (jj[q] < TS) {
                        const cv<R> v = tile[lds_pos(jj[q] ^ xl)];
                        vr[q] = (double)v.x;
                        vi[q] = (double)v.y;
                    }
                }
                const int k0 = gp[g].first, km = k0 + gp[g].nre, k1 = k0 + gp[g].count;
                const int kr = k0 + gp[g].nre0, ki = km + gp[g].nim0;
                double cr0 = 0.0, ci0 = 0.0;  // the terms without sign bits on the tile
                for (int k = k0; k < kr; ++k) cr0 += csg[k];
                for (int k = km; k < ki; ++k) ci0 += csg[k];
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    cr[q] = cr0;
                    ci[q] = ci0;
                }
                uint32_t jx[E];
#pragma unroll
                for (int q = 0; q < E; ++q) jx[q] = jj[q] ^ xl;
                // one run of terms into acc: each term is the same for every
                // lane (scalar registers); its coefficient's sign bit is
                // flipped by the partner's parity
                auto run = [&](int ka, int kb, double(&acc)[E]) __attribute__((always_inline)) {
#pragma unroll 2
                    for (int k = ka; k < kb; ++k) {
                        const double c = csg[k];
                        const uint32_t zl = __builtin_amdgcn_readfirstlane(zls[k]);
                        const int chi = __builtin_amdgcn_readfirstlane(__double2hiint(c)),
                                  clo = __builtin_amdgcn_readfirstlane(__double2loint(c));
#pragma unroll
                        for (int q = 0; q < E; ++q) {
                            const int hi = chi ^ (int)((__popc(jx[q] & zl) & 1u) << 31);
                            acc[q] += __hiloint2double(hi, clo);
                        }
                    }
                };
                run(kr, km, cr);
                run(ki, k1, ci);
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    sr[q] += cr[q] * vr[q] - ci[q] * vi[q];
                    si[q] += cr[q] * vi[q] + ci[q] * vr[q];
                }
            }
#pragma unroll
            for (int q = 0; q < E; ++q) {
                if (jj[q] >= TS) continue;
                const cv<R> p = tile[lds_pos(jj[q])];
                e += (double)p.x * sr[q] + (double)p.y * si[q];
                if (WRITE) {
                    const uint64_t at = pass_global(d, t, jj[q]);
                    cv<R> o;
                    if (FIRST) {
                        o.x = (R)sr[q];
                        o.y = (R)si[q];
                    } else {
                        o = lam[at];
                        o.x = (R)((double)o.x + sr[q]);
                        o.y = (R)((double)o.y + si[q]);
                    }
                    lam[at] = o;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
    if ((tid & 63) == 0) wsum[tid >> 6] = e;
    __syncthreads();
    if (tid == 0) {
        double s2 = 0.0;
        for (uint32_t w = 0; w < (T >> 6); ++w) s2 += wsum[w];
        part[blockIdx.x] = s2;
    }
}

// Energy only: Re sum_i psi_i * S_i with S_i = sum_g f_g(i) conj(psi[i ^ x_g])
// and f_g(i) = sum_t c_t i^{nY} (-1)^{popcount(i & zy_t)}: the reference's
// sum over i (core/src/statevector.cpp:265-288) per term.  Summing on the
// reference's side puts each phase on the element's own index and SKIP
// (the reference's |psi_i|^2 < 1e-18 test) on the element itself, once per
// element instead of once per group; a group whose terms have no sign bits on
// the tile has one phase sum per tile (every single-qubit X of a field term).
template <typename R, bool SKIP>
__global__ void __launch_bounds__(512) k_obs_energy_tile(const cv<R>* __restrict__ psi, ObsPass op,
                                                         const ObsGroupT* __restrict__ groups,
                                                         const ObsTermT* __restrict__ terms,
                                                         const ObsChainT* __restrict__ chains, double* __restrict__ part) {
    constexpr int E = 4;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double wsum[16];
    const PassDesc& d = op.d;
    const uint32_t T = blockDim.x, tid = threadIdx.x, TS = 1u << d.ts_bits;
    cv<R>* tile = reinterpret_cast<cv<R>*>(smem);
    double* csg = reinterpret_cast<double*>(tile + TS);
    uint32_t* zls = reinterpret_cast<uint32_t*>(csg + op.tcount);
    const ObsGroupT* gp = groups + op.gfirst;
    const ObsTermT* tp = terms + op.tfirst;
    for (int k = tid; k < op.tcount; k += T) zls[k] = tp[k].zl;
    double e = 0.0, ei = 0.0;  // Re and (SKIP) Im of sum_i psi_i S_i
    for (uint64_t t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        tile_in<R>(d, t, psi, tile, TS);
        const uint64_t base = pass_global(d, t, 0);
        for (int k = tid; k < op.tcount; k += T) csg[k] = (__popcll(base & tp[k].zfix) & 1) ? -tp[k].coeff : tp[k].coeff;
        __syncthreads();
        for (uint32_t c0 = 0; c0 < TS; c0 += E * T) {
            uint32_t jj[E];
            double pr[E], pi[E], sr[E], si[E];
#pragma unroll
            for (int q = 0; q < E; ++q) {
                jj[q] = c0 + q * T + tid;
                pr[q] = pi[q] = sr[q] = si[q] = 0.0;
                if (jj[q] < TS) {
                    const cv<R> v = tile[lds_pos(jj[q])];
                    pr[q] = (double)v.x;
                    pi[q] = (double)v.y;
                    if (SKIP && pr[q] * pr[q] + pi[q] * pi[q] < 1e-18) pr[q] = pi[q] = 0.0;
                }
            }
            for (int gi = 0; gi < op.gcount; ++gi) {
                const int g = __builtin_amdgcn_readfirstlane(gi);
                const uint32_t xl = gp[g].xl;
                const int k0 = gp[g].first, km = k0 + gp[g].nre, k1 = k0 + gp[g].count;
                double wr[E], wi[E];
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    const cv<R> w = tile[lds_pos((jj[q] ^ xl) & (TS - 1))];
                    wr[q] = (double)w.x;
                    wi[q] = (double)w.y;
                }
                // the terms without sign bits on the tile: one sum per tile
                const int kr = k0 + gp[g].nre0, ki = km + gp[g].nim0;
                double fr0 = 0.0, fi0 = 0.0;
                for (int k = k0; k < kr; ++k) fr0 += csg[k];
                for (int k = km; k < ki; ++k) fi0 += csg[k];
                if (kr == km && ki == k1) {
                    if (km == k1) {
#pragma unroll
                        for (int q = 0; q < E; ++q) {
                            sr[q] += fr0 * wr[q];
                            si[q] -= fr0 * wi[q];
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < E; ++q) {
                            sr[q] += fr0 * wr[q] + fi0 * wi[q];
                            si[q] += fi0 * wr[q] - fr0 * wi[q];
                        }
                    }
                    continue;
                }
                double fr[E], fi[E];
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    fr[q] = fr0;
                    fi[q] = fi0;
                }
                // chain classes: coeff * (count - 2 popcount(f(j) & mask))
                const int c0 = gp[g].cfirst, c1 = c0 + gp[g].ccount;
                for (int ci = c0; ci < c1; ++ci) {
                    const int c = __builtin_amdgcn_readfirstlane(ci);
                    const uint32_t mask = chains[c].mask;
                    const int sh = chains[c].shift, cnt = chains[c].count;
                    const double cc = chains[c].coeff;
#pragma unroll
                    for (int q = 0; q < E; ++q) {
                        const uint32_t f = sh ? (jj[q] ^ (jj[q] >> sh)) : jj[q];
                        fr[q] += cc * (double)(cnt - 2 * (int)__popc(f & mask));
                    }
                }
                auto run = [&](int ka, int kb, double(&acc)[E]) __attribute__((always_inline)) {
                    for (int k = ka; k < kb; ++k) {
                        const double c = csg[k];
                        const uint32_t zl = __builtin_amdgcn_readfirstlane(zls[k]);
                        const int chi = __builtin_amdgcn_readfirstlane(__double2hiint(c)),
                                  clo = __builtin_amdgcn_readfirstlane(__double2loint(c));
#pragma unroll
                        for (int q = 0; q < E; ++q) {
                            const int hi = chi ^ (int)((__popc(jj[q] & zl) & 1u) << 31);
                            acc[q] += __hiloint2double(hi, clo);
                        }
                    }
                };
                run(kr + gp[g].nch, km, fr);
                run(ki, k1, fi);
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    sr[q] += fr[q] * wr[q] + fi[q] * wi[q];
                    si[q] += fi[q] * wr[q] - fr[q] * wi[q];
                }
            }
#pragma unroll
            for (int q = 0; q < E; ++q) {
                e += pr[q] * sr[q] - pi[q] * si[q];
                if (SKIP) ei += pr[q] * si[q] + pi[q] * sr[q];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_down(e, off, 64);
        if (SKIP) ei += __shfl_down(ei, off, 64);
    }
    if ((tid & 63) == 0) {
        wsum[tid >> 6] = e;
        wsum[8 + (tid >> 6)] = ei;
    }
    __syncthreads();
    if (tid == 0) {
        double s2 = 0.0, s3 = 0.0;
        for (uint32_t w = 0; w < (T >> 6); ++w) {
            s2 += wsum[w];
            s3 += wsum[8 + w];
        }
        if (SKIP) {
            part[2 * blockIdx.x] = s2;
            part[2 * blockIdx.x + 1] = s3;
        } else {
            part[blockIdx.x] = s2;
        }
    }
}

__global__ void __launch_bounds__(256) k_finalize_slots(const double* __restrict__ part, int nblocks, int nslots,
                                                        double* __restrict__ out) {
    // one block per output: slot = blockIdx.x / kSubMax, sub-step = blockIdx.x % kSubMax
    const int slot = blockIdx.x / kSubMax, s = blockIdx.x % kSubMax;
    if (slot >= nslots) return;
    __shared__ double sh[4];
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v += part[((uint64_t)slot * nblocks + b) * kSubMax + s];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

inline unsigned nblocks_for(uint64_t items) {
    uint64_t g = (items + kRedThreads - 1) / kRedThreads;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, kRedBlocks));
}

using cplx = std::complex<double>;
struct M2 {
    cplx m[4];
};

M2 mat_rx(double t) { const double c = std::cos(t / 2), s = std::sin(t / 2); return {{c, cplx(0, -s), cplx(0, -s), c}}; }
M2 mat_ry(double t) { const double c = std::cos(t / 2), s = std::sin(t / 2); return {{c, -s, s, c}}; }
M2 mat_rz(double t) { return {{std::exp(cplx(0, -t / 2)), 0.0, 0.0, std::exp(cplx(0, t / 2))}}; }
M2 mul(const M2& a, const M2& b) {  // a * b
    return {{a.m[0] * b.m[0] + a.m[1] * b.m[2], a.m[0] * b.m[1] + a.m[1] * b.m[3],
             a.m[2] * b.m[0] + a.m[3] * b.m[2], a.m[2] * b.m[1] + a.m[3] * b.m[3]}};
}
M2 dag(const M2& a) { return {{std::conj(a.m[0]), std::conj(a.m[2]), std::conj(a.m[1]), std::conj(a.m[3])}}; }
M2 pauli(int gen) {
    switch (gen) {
        case kGenX: return {{0.0, 1.0, 1.0, 0.0}};
        case kGenY: return {{0.0, cplx(0, -1), cplx(0, 1), 0.0}};
        case kGenZ: return {{1.0, 0.0, 0.0, -1.0}};
        default: return {{0.0, 0.0, 0.0, 0.0}};
    }
}

}  // namespace

struct TapeGate {
    int kind = 0;
    int q0 = 0, q1 = 0;
    int64_t p[3] = {-1, -1, -1};
    double mre[16] = {}, mim[16] = {};
};

struct Sub {
    int gate;
    int which;  // U3 sub-step: 0 = RZ(lam), 1 = RY(theta), 2 = RZ(phi)
};

struct Segment {
    int kind = 0;  // 0 = 1q, 1 = diag, 2 = 2q
    int q = -1;
    std::vector<Sub> subs;
    std::vector<std::pair<int, int>> edges;
    int gate = -1;
    int slot = -1;
};

// A Pauli sum over tile passes: launches over the passes whose tile bits hold
// a group's flip mask, then the groups no pass holds through global gathers;
// energy partials of every launch in opart.
struct ObsPlan {
    std::vector<ObsPass> opass;
    void* otab_dev = nullptr;
    const ObsGroupT* ogroups = nullptr;
    const ObsTermT* oterms = nullptr;
    const ObsChainT* ochains = nullptr;
    ObsTable far{};
    void* far_dev = nullptr;
    double* opart = nullptr;
    size_t opart_len = 0;
    bool obs_tiled = false;

    ObsPlan() = default;
    ObsPlan(const ObsPlan&) = delete;
    ObsPlan& operator=(const ObsPlan&) = delete;
    void release() {
        if (otab_dev) hipFree(otab_dev);
        if (far_dev) hipFree(far_dev);
        if (opart) hipFree(opart);
        otab_dev = far_dev = nullptr;
        opart = nullptr;
        ogroups = nullptr;
        oterms = nullptr;
        ochains = nullptr;
        far = ObsTable{};
        opass.clear();
        opart_len = 0;
        obs_tiled = false;
    }
    ~ObsPlan() { release(); }
};

struct Circuit {
    int n = 0, prec = kC128;
    Engine eng;
    void* lam = nullptr;
    std::vector<TapeGate> gates;
    std::vector<Segment> segs;
    bool compiled = false;
    int nslots = 0;
    unsigned nblk = 1;
    double* gpart = nullptr;
    double* gout = nullptr;
    double* hgout = nullptr;
    size_t gcap = 0;
    // observable
    std::vector<double> coeff;
    std::vector<PauliTerm> terms;
    void* obs_dev = nullptr;
    ObsTable obs{};
    double* evals = nullptr;  // device energies for parameter shift
    size_t evals_cap = 0;
    // tiled stages: pass descriptors go up through a small ring of pinned
    // (coherent) + device buffers, so consecutive evaluations need no sync
    struct DescSlot {
        void* host = nullptr;
        void* host_dev = nullptr;
        void* dev = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        bool pending = false;
    };
    DescSlot dring[4];
    int dnext = 0;
    // Stored states of a tiled forward sweep: pass i writes stored[i] (the
    // last pass writes the engine state), and the reverse sweep reads them
    // instead of taking psi back.  Used while they fit stored_budget bytes.
    void* stored = nullptr;
    size_t stored_cap = 0;   // bytes
    size_t stored_budget = size_t(16) << 30;
    bool stored_valid = false;  // the last forward sweep filled them
    // Reverse passes over stored outputs stage lambda alone in LDS and read
    // psi from its slot (true), or stage both (false); the same sums either way.
    bool rev_lam_only = true;
    ObsPlan oplan;  // the observable over tile passes (16 to 34 qubits)
    // The forward sweep's plan (pass descriptors, launches, phase tables) of a
    // tape with no parametric gate, which depends on the tape alone: made on
    // the first run, dropped when a gate is added.
    std::shared_ptr<void> fixed_plan;
    // Per-launch timing (off by default): each launch of an evaluation is
    // bracketed by two events on the circuit's stream; kinds are kLaunch*.
    bool timing = false;
    std::vector<hipEvent_t> tev;  // start, end pairs
    std::vector<int> tkind;
    size_t tused = 0;

    ~Circuit() {
        if (eng.stream) hipStreamSynchronize(eng.stream);
        for (hipEvent_t ev : tev) hipEventDestroy(ev);
        if (lam) hipFree(lam);
        if (stored) hipFree(stored);
        if (gpart) hipFree(gpart);
        if (gout) hipFree(gout);
        if (hgout) hipHostFree(hgout);
        if (obs_dev) hipFree(obs_dev);
        if (evals) hipFree(evals);
        for (auto& d : dring) {
            if (d.host) hipHostFree(d.host);
            if (d.dev) hipFree(d.dev);
            if (d.ev) hipEventDestroy(d.ev);
        }
    }
};

namespace {

enum LaunchKind : int { kLaunchForward = 0, kLaunchObservable = 1, kLaunchReverse = 2, kLaunchTwoQubit = 3,
                        kLaunchFinalize = 4 };

// Brackets the launches made during its lifetime with events when timing is on.
struct TimedLaunch {
    Circuit& c;
    size_t slot = SIZE_MAX;
    TimedLaunch(Circuit& circ, int kind) : c(circ) {
        if (!c.timing) return;
        if (c.tused == c.tkind.size()) {
            hipEvent_t a = nullptr, b = nullptr;
            if (hipEventCreate(&a) != hipSuccess) return;
            if (hipEventCreate(&b) != hipSuccess) {
                hipEventDestroy(a);
                return;
            }
            c.tev.push_back(a);
            c.tev.push_back(b);
            c.tkind.push_back(kind);
        }
        slot = c.tused++;
        c.tkind[slot] = kind;
        hipEventRecord(c.tev[2 * slot], c.eng.stream);
    }
    ~TimedLaunch() {
        if (slot != SIZE_MAX) hipEventRecord(c.tev[2 * slot + 1], c.eng.stream);
    }
};

M2 sub_matrix(const TapeGate& g, int which, const double* params) {
    auto P = [&](int i) { return params[g.p[i]]; };
    switch (g.kind) {
        case kRX: return mat_rx(P(0));
        case kRY: return mat_ry(P(0));
        case kRZ: return mat_rz(P(0));
        case kU3: return which == 0 ? mat_rz(P(2)) : (which == 1 ? mat_ry(P(0)) : mat_rz(P(1)));
        default: {
            M2 m;
            for (int i = 0; i < 4; ++i) m.m[i] = cplx(g.mre[i], g.mim[i]);
            return m;
        }
    }
}

int sub_gen(const TapeGate& g, int which) {
    switch (g.kind) {
        case kRX: return kGenX;
        case kRY: return kGenY;
        case kRZ: return kGenZ;
        case kU3: return which == 1 ? kGenY : kGenZ;
        default: return kGenNone;
    }
}

int64_t sub_param(const TapeGate& g, int which) {
    switch (g.kind) {
        case kRX: case kRY: case kRZ: return g.p[0];
        case kU3: return which == 0 ? g.p[2] : (which == 1 ? g.p[0] : g.p[1]);
        default: return -1;
    }
}

bool is_1q(int kind) { return kind == kRX || kind == kRY || kind == kRZ || kind == kU3 || kind == kMat1; }

// Dependency-respecting reordering: ops sharing a qubit keep their order,
// except CZ-CZ pairs, which commute.  Greedy list scheduling that emits runs of
// one-qubit gates per qubit, then every ready CZ as one sign pass.
void compile(Circuit& c) {
    const int L = (int)c.gates.size();
    std::vector<std::vector<int>> preds(L), succs(L);
    std::vector<int> last_nondiag(c.n, -1);
    std::vector<std::vector<int>> diag_since(c.n);
    for (int i = 0; i < L; ++i) {
        const TapeGate& g = c.gates[i];
        const bool diag = g.kind == kCZ;
        std::vector<int> qs = {g.q0};
        if (!is_1q(g.kind)) qs.push_back(g.q1);
        for (int q : qs) {
            if (last_nondiag[q] >= 0) preds[i].push_back(last_nondiag[q]);
            if (diag) {
                diag_since[q].push_back(i);
            } else {
                for (int d : diag_since[q]) preds[i].push_back(d);
                diag_since[q].clear();
                last_nondiag[q] = i;
            }
        }
        std::sort(preds[i].begin(), preds[i].end());
        preds[i].erase(std::unique(preds[i].begin(), preds[i].end()), preds[i].end());
        for (int p : preds[i]) succs[p].push_back(i);
    }
    std::vector<int> indeg(L), done(L, 0);
    for (int i = 0; i < L; ++i) indeg[i] = (int)preds[i].size();
    auto ready = [&](int i) { return !done[i] && indeg[i] == 0; };
    auto finish = [&](int i) {
        done[i] = 1;
        for (int s : succs[i]) --indeg[s];
    };
    c.segs.clear();
    int emitted = 0;
    while (emitted < L) {
        bool progress = false;
        // runs of one-qubit gates, qubit by qubit
        for (int q = 0; q < c.n; ++q) {
            Segment seg;
            seg.kind = 0;
            seg.q = q;
            for (;;) {
                int pick = -1;
                for (int i = 0; i < L; ++i)
                    if (ready(i) && is_1q(c.gates[i].kind) && c.gates[i].q0 == q) { pick = i; break; }
                if (pick < 0) break;
                const int nsub = c.gates[pick].kind == kU3 ? 3 : 1;
                if ((int)seg.subs.size() + nsub > kSubMax) {
                    c.segs.push_back(seg);
                    seg.subs.clear();
                }
                for (int w = 0; w < nsub; ++w) seg.subs.push_back({pick, w});
                finish(pick);
                ++emitted;
            }
            if (!seg.subs.empty()) {
                c.segs.push_back(seg);
                progress = true;
            }
        }
        // every ready CZ as one sign pass
        Segment d;
        d.kind = 1;
        for (int i = 0; i < L; ++i)
            if (ready(i) && c.gates[i].kind == kCZ) {
                if ((int)d.edges.size() == kEdgeMax) {
                    c.segs.push_back(d);
                    d.edges.clear();
                }
                d.edges.push_back({c.gates[i].q0, c.gates[i].q1});
                finish(i);
                ++emitted;
            }
        if (!d.edges.empty()) {
            c.segs.push_back(d);
            progress = true;
        }
        if (progress) continue;
        for (int i = 0; i < L; ++i)
            if (ready(i)) {
                Segment s;
                s.kind = 2;
                s.gate = i;
                c.segs.push_back(s);
                finish(i);
                ++emitted;
                break;
            }
    }
    c.nslots = 0;
    for (auto& s : c.segs)
        if (s.kind == 0) s.slot = c.nslots++;
    c.compiled = true;
}

bool ensure_grad_buffers(Circuit& c) {
    const size_t need = (size_t)std::max(c.nslots, 1);
    if (c.gcap >= need) return true;
    if (c.gpart) hipFree(c.gpart);
    if (c.gout) hipFree(c.gout);
    if (c.hgout) hipHostFree(c.hgout);
    c.gpart = c.gout = c.hgout = nullptr;
    c.gcap = 0;
    if (!hip_ok(hipMalloc(&c.gpart, need * c.nblk * kSubMax * sizeof(double)), "gradient partials") ||
        !hip_ok(hipMalloc(&c.gout, need * kSubMax * sizeof(double)), "gradient slots") ||
        !hip_ok(hipHostMalloc((void**)&c.hgout, need * kSubMax * sizeof(double)), "gradient host slots"))
        return false;
    c.gcap = need;
    return true;
}

bool check_params(const Circuit& c, size_t nparams) {
    for (const auto& g : c.gates)
        for (int i = 0; i < 3; ++i)
            if (g.p[i] >= 0 && (size_t)g.p[i] >= nparams) {
                set_error(kErrInvalidArgument, "gate uses parameter %lld but only %zu were given", (long long)g.p[i], nparams);
                return false;
            }
    return true;
}

template <typename R>
bool run_forward_segments(Circuit& c, const double* params) {
    Engine& e = c.eng;
    if (!e.reset_basis(0)) return false;
    for (const auto& s : c.segs) {
        if (s.kind == 0) {
            M2 m = {{1.0, 0.0, 0.0, 1.0}};
            for (const auto& sub : s.subs) m = mul(sub_matrix(c.gates[sub.gate], sub.which, params), m);
            double re[4], im[4];
            for (int i = 0; i < 4; ++i) {
                re[i] = m.m[i].real();
                im[i] = m.m[i].imag();
            }
            GateDesc g;
            g.K = 1;
            g.q[0] = (uint32_t)s.q;
            g.re = re;
            g.im = im;
            if (!hip_ok(launch_gate(e.amp, c.n, c.prec, e.stream, g), "forward 1q")) return false;
        } else if (s.kind == 1) {
            EdgeList el{};
            el.ne = (int)s.edges.size();
            for (int k = 0; k < el.ne; ++k) {
                el.a[k] = (uint32_t)s.edges[k].first;
                el.b[k] = (uint32_t)s.edges[k].second;
            }
            const uint64_t count = e.dim;
            const unsigned b = (unsigned)std::min<uint64_t>((count + 255) / 256, 65536);
            hipLaunchKernelGGL(k_cz_set<R>, dim3(b), dim3(256), 0, e.stream, (cv<R>*)e.amp, (cv<R>*)nullptr, count, el);
            if (!hip_ok(hipGetLastError(), "forward CZ")) return false;
        } else {
            const TapeGate& tg = c.gates[s.gate];
            double re[16], im[16];
            if (tg.kind == kCX) {
                const double cx[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0};
                std::memcpy(re, cx, sizeof(re));
                std::memset(im, 0, sizeof(im));
            } else {
                std::memcpy(re, tg.mre, sizeof(re));
                std::memcpy(im, tg.mim, sizeof(im));
            }
            GateDesc g;
            g.K = 2;
            g.q[0] = (uint32_t)tg.q0;
            g.q[1] = (uint32_t)tg.q1;
            g.re = re;
            g.im = im;
            if (!hip_ok(launch_gate(e.amp, c.n, c.prec, e.stream, g), "forward 2q")) return false;
        }
    }
    return true;
}

// The planned Pauli sum over the state `amp` (dim amplitudes): Re <psi|H psi>
// into energy_dev and, with `write`, H psi into lam.  `skip` (energy only)
// applies the reference's |psi_i|^2 < 1e-18 skip and writes Re and Im into
// energy_dev[0], energy_dev[1] (the plan holds two partials per block).
template <typename R>
bool launch_observable_plan(const ObsPlan& p, const void* amp, void* lam, uint64_t dim, hipStream_t st, bool write,
                            bool skip, double* energy_dev) {
    if (write && skip) return false;
    const int m = skip ? 2 : 1;  // partials per block
    size_t off = 0;
    bool first = true;
    for (const ObsPass& op : p.opass) {
        const unsigned grid = (unsigned)std::min<uint64_t>(op.d.ntiles, 2048);
        const size_t lds = ((size_t)1 << op.d.ts_bits) * sizeof(cv<R>) + (size_t)op.tcount * (sizeof(double) + 4);
        if (!write) {
            auto kern = skip ? k_obs_energy_tile<R, true> : k_obs_energy_tile<R, false>;
            hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(512), lds, st, (const cv<R>*)amp, op, p.ogroups, p.oterms,
                               p.ochains, p.opart + m * off);
        } else {
            auto kern = first ? k_obs_tile<R, true, true> : k_obs_tile<R, true, false>;
            hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(512), lds, st, (const cv<R>*)amp, (cv<R>*)lam, op, p.ogroups,
                               p.oterms, p.opart + off);
        }
        if (!hip_ok(hipGetLastError(), "tiled observable")) return false;
        off += grid;
        first = false;
    }
    if (p.far.ngroups > 0) {
        const unsigned b = nblocks_for(dim);
        auto kern = skip ? k_apply_h<R, false, false, true>
                         : !write ? k_apply_h<R, false> : (first ? k_apply_h<R, true, false> : k_apply_h<R, true, true>);
        hipLaunchKernelGGL(kern, dim3(b), dim3(256), 0, st, (const cv<R>*)amp, (cv<R>*)lam, dim, p.far, p.opart + m * off);
        if (!hip_ok(hipGetLastError(), "observable")) return false;
        off += b;
    }
    return hip_ok(launch_finalize(p.opart, (unsigned)off, m, energy_dev, st), "finalize");
}

template <typename R>
bool apply_observable_tiled(Circuit& c, bool write, double* energy_dev) {
    return launch_observable_plan<R>(c.oplan, c.eng.amp, c.lam, c.eng.dim, c.eng.stream, write, false, energy_dev);
}

template <typename R>
bool apply_observable(Circuit& c, bool write, double* energy_dev) {
    Engine& e = c.eng;
    TimedLaunch tl(c, kLaunchObservable);
    if (c.oplan.obs_tiled) return apply_observable_tiled<R>(c, write, energy_dev);
    const unsigned b = nblocks_for(e.dim);
    if (write)
        hipLaunchKernelGGL((k_apply_h<R, true>), dim3(b), dim3(256), 0, e.stream, (const cv<R>*)e.amp, (cv<R>*)c.lam, e.dim,
                           c.obs, e.red_part);
    else
        hipLaunchKernelGGL((k_apply_h<R, false>), dim3(b), dim3(256), 0, e.stream, (const cv<R>*)e.amp, (cv<R>*)nullptr,
                           e.dim, c.obs, e.red_part);
    return hip_ok(hipGetLastError(), "observable") && hip_ok(launch_finalize(e.red_part, b, 1, energy_dev, e.stream), "finalize");
}

template <typename R>
bool run_reverse_segments(Circuit& c, const double* params) {
    Engine& e = c.eng;
    const uint64_t npairs = e.dim >> 1;
    for (int si = (int)c.segs.size() - 1; si >= 0; --si) {
        const Segment& s = c.segs[si];
        if (s.kind == 0) {
            Seg1qArgs a{};
            a.count = (int)s.subs.size();
            for (int k = 0; k < a.count; ++k) {
                const TapeGate& g = c.gates[s.subs[k].gate];
                const M2 d = dag(sub_matrix(g, s.subs[k].which, params));
                for (int i = 0; i < 4; ++i) {
                    a.u[k][2 * i] = d.m[i].real();
                    a.u[k][2 * i + 1] = d.m[i].imag();
                }
                a.gen[k] = sub_gen(g, s.subs[k].which);
            }
            hipLaunchKernelGGL(k_adj_seg1q<R>, dim3(c.nblk), dim3(256), 0, e.stream, (cv<R>*)e.amp, (cv<R>*)c.lam, npairs,
                               s.q, a, c.gpart + (size_t)s.slot * c.nblk * kSubMax);
            if (!hip_ok(hipGetLastError(), "reverse 1q")) return false;
        } else if (s.kind == 1) {
            EdgeList el{};
            el.ne = (int)s.edges.size();
            for (int k = 0; k < el.ne; ++k) {
                el.a[k] = (uint32_t)s.edges[k].first;
                el.b[k] = (uint32_t)s.edges[k].second;
            }
            const unsigned b = (unsigned)std::min<uint64_t>((e.dim + 255) / 256, 65536);
            hipLaunchKernelGGL(k_cz_set<R>, dim3(b), dim3(256), 0, e.stream, (cv<R>*)e.amp, (cv<R>*)c.lam, e.dim, el);
            if (!hip_ok(hipGetLastError(), "reverse CZ")) return false;
        } else {
            const TapeGate& tg = c.gates[s.gate];
            double re[16], im[16];
            if (tg.kind == kCX) {  // CX is its own inverse
                const double cx[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0};
                std::memcpy(re, cx, sizeof(re));
                std::memset(im, 0, sizeof(im));
            } else {
                for (int r = 0; r < 4; ++r)
                    for (int col = 0; col < 4; ++col) {
                        re[r * 4 + col] = tg.mre[col * 4 + r];
                        im[r * 4 + col] = -tg.mim[col * 4 + r];
                    }
            }
            GateDesc g;
            g.K = 2;
            g.q[0] = (uint32_t)tg.q0;
            g.q[1] = (uint32_t)tg.q1;
            g.re = re;
            g.im = im;
            if (!hip_ok(launch_gate(e.amp, c.n, c.prec, e.stream, g), "reverse 2q psi") ||
                !hip_ok(launch_gate(c.lam, c.n, c.prec, e.stream, g), "reverse 2q lambda"))
                return false;
        }
    }
    if (c.nslots > 0) {
        hipLaunchKernelGGL(k_finalize_slots, dim3(c.nslots * kSubMax), dim3(256), 0, e.stream, c.gpart, (int)c.nblk,
                           c.nslots, c.gout);
        if (!hip_ok(hipGetLastError(), "gradient finalize")) return false;
    }
    return true;
}

// ----------------------------------------------------------- tiled stages
// Stage plan: consecutive one-qubit runs (any qubits, order kept per qubit)
// closed by a CZ set; a two-qubit gate stands alone between stages.
struct TileGeom {
    int b, r;  // pass-0 tile bits; run bits of later passes
};

// 64 KiB tiles: with 512 threads and register groups of up to 4 (complex64)
// or 3 (complex128) qubits, every thread owns one sub-cube per group.  Runs
// of 2^r amplitudes are 128-byte lines.
TileGeom tile_geom(int prec) { return prec == kC64 ? TileGeom{13, 4} : TileGeom{12, 3}; }
TileGeom tile_geom(const Circuit& c) { return tile_geom(c.prec); }

// tiled stages pay off once the state no longer fits a few tiles
bool use_tiles(const Circuit& c) { return c.n >= 16 && c.n <= 34; }

// Qubits [b, n) are split evenly over the passes after pass 0, at most b - r
// each; a pass's run bits fill the rest of its 2^b tile (>= r of them), so no
// pass degenerates into many tiny tiles.
int high_passes(const TileGeom& g, int n) { return n <= g.b ? 0 : (n - g.b + (g.b - g.r) - 1) / (g.b - g.r); }

PassDesc pass_desc(const TileGeom& g, int n, int k) {
    PassDesc d{};
    if (k == 0) {
        d.r = std::min(g.b, n);
        d.m = 0;
        d.g0 = d.r;
    } else {
        const int P = high_passes(g, n), total = n - g.b, base = total / P, extra = total % P;
        const int kk = P + 1 - k;  // pass 1 takes the highest qubits
        d.m = base + (kk - 1 < extra ? 1 : 0);
        d.g0 = g.b + (kk - 1) * base + std::min(kk - 1, extra);
        d.r = g.b - d.m;
    }
    d.ts_bits = d.r + d.m;
    d.ntiles = 1ULL << (n - d.ts_bits);
    d.greg = d.tpre = d.tpost = -1;
    return d;
}

int pass_of(const TileGeom& g, int n, int q) {
    if (q < g.b) return 0;
    for (int k = 1;; ++k) {
        const PassDesc d = pass_desc(g, n, k);
        if (q >= d.g0 && q < d.g0 + d.m) return k;
    }
}

int tile_bit(const TileGeom&, const PassDesc& d, int k, int q) { return k == 0 ? q : d.r + (q - d.g0); }

// Tile bit of global qubit q in pass d, or -1 when q is fixed by the tile index.
int local_bit(const PassDesc& d, int q) {
    if (q < d.r) return q;
    if (d.m > 0 && q >= d.g0 && q < d.g0 + d.m) return d.r + (q - d.g0);
    return -1;
}

// Files a CZ set into the pass descriptor by where each edge's qubits sit
// (the split cz_odd evaluates).  A repeated edge cancels: masks combine by XOR.
void set_cz(PassDesc& d, const std::vector<std::pair<int, int>>& edges) {
    d.ncz = (int)edges.size();
    d.ncross = d.nfixed = d.nshift = 0;
    // cz_odd evaluates the first two distance classes unconditionally: unused
    // entries must contribute nothing
    for (int k = 0; k < kShiftMax; ++k) d.qshift[k] = d.qmask[k] = 0;
    std::vector<std::pair<int, int>> fixed;
    for (const auto& e : edges) {
        int la = local_bit(d, e.first), lb = local_bit(d, e.second);
        if (la >= 0 && lb >= 0) {
            if (la > lb) std::swap(la, lb);
            const int sh = lb - la;
            int k = 0;
            while (k < d.nshift && d.qshift[k] != sh) ++k;
            if (k == d.nshift) {
                d.qshift[k] = (uint32_t)sh;
                d.qmask[k] = 0;
                ++d.nshift;
            }
            d.qmask[k] ^= 1u << la;
        } else if (la >= 0 || lb >= 0) {
            d.cza[d.ncross] = (uint32_t)(la >= 0 ? la : lb);
            d.czb[d.ncross] = (uint32_t)(la >= 0 ? e.second : e.first);
            ++d.ncross;
        } else {
            fixed.push_back(e);
        }
    }
    for (const auto& e : fixed) {
        d.cza[d.ncross + d.nfixed] = (uint32_t)e.first;
        d.czb[d.ncross + d.nfixed] = (uint32_t)e.second;
        ++d.nfixed;
    }
}

struct Launch {
    bool tile = true;
    PassDesc d{};
    int seg2 = -1;  // a two-qubit segment
};

// Passes of the whole tape in forward order (reverse = false) or the reverse
// sweep's order, with their run descriptors.
void plan_tiles(const Circuit& c, const double* params, bool reverse, std::vector<PassSeg>& segv,
                std::vector<Launch>& launches, std::vector<double>& tabs) {
    const TileGeom g = tile_geom(c);
    struct Stage {
        std::vector<int> runs;
        int cz = -1, seg2 = -1;
    };
    std::vector<Stage> stages;
    Stage cur;
    for (int i = 0; i < (int)c.segs.size(); ++i) {
        const Segment& sg = c.segs[i];
        if (sg.kind == 0) {
            cur.runs.push_back(i);
        } else if (sg.kind == 1) {
            cur.cz = i;
            stages.push_back(cur);
            cur = Stage{};
        } else {
            if (!cur.runs.empty()) stages.push_back(cur);
            cur = Stage{};
            Stage two;
            two.seg2 = i;
            stages.push_back(two);
        }
    }
    if (!cur.runs.empty() || cur.cz >= 0) stages.push_back(cur);
    const int npass = high_passes(g, c.n) + 1;
    auto fill = [&](const Segment& sg, const PassDesc& d, int k) {
        PassSeg ps{};
        ps.tbit = tile_bit(g, d, k, sg.q);
        ps.slot = sg.slot;
        // the run's product (forward) or its inverse (reverse sweep)
        M2 m = {{1.0, 0.0, 0.0, 1.0}};
        for (const auto& sub : sg.subs) m = mul(sub_matrix(c.gates[sub.gate], sub.which, params), m);
        if (reverse) m = dag(m);
        for (int i = 0; i < 4; ++i) {
            ps.u[2 * i] = ps.v[2 * i] = m.m[i].real();
            ps.u[2 * i + 1] = ps.v[2 * i + 1] = m.m[i].imag();
        }
        segv.push_back(ps);
    };
    std::vector<Launch> fwd;  // in forward order; reversed at the end for the sweep
    std::vector<std::vector<int>> run_chunks;  // per forward launch: the runs, in forward order
    for (const Stage& st : stages) {
        if (st.seg2 >= 0) {
            Launch l;
            l.tile = false;
            l.seg2 = st.seg2;
            fwd.push_back(l);
            run_chunks.push_back({});
            continue;
        }
        int last = -1;  // index in fwd of the stage's last pass
        const int gmax = c.prec == kC64 ? kGroupRuns : 3;
        for (int k = 0; k < npass; ++k) {
            const PassDesc dk = pass_desc(g, c.n, k);
            std::vector<int> mine;
            for (int ri : st.runs)
                if (pass_of(g, c.n, c.segs[ri].q) == k) mine.push_back(ri);
            if (mine.empty()) continue;
            // runs on different qubits commute: sort by tile bit (stable, so a
            // qubit's own runs keep their order); cut the chains of consecutive
            // tile bits (one run each) into groups of at most gmax, of about
            // equal size (a lone short group costs a whole LDS round trip)
            std::stable_sort(mine.begin(), mine.end(), [&](int a, int b2) {
                return tile_bit(g, dk, k, c.segs[a].q) < tile_bit(g, dk, k, c.segs[b2].q);
            });
            auto tb_of = [&](int ri) { return tile_bit(g, dk, k, c.segs[ri].q); };
            std::vector<std::pair<size_t, int>> groups;  // first entry of mine, size
            // runs on exactly tile bits 9 .. 8 + gmax (one each, the pass's
            // highest) form the register group, applied as the tile is stored
            size_t nchain = mine.size();
            bool reg = dk.ts_bits == 9 + gmax && mine.size() >= (size_t)gmax;
            for (int q = 0; reg && q < gmax; ++q)
                reg = tb_of(mine[mine.size() - gmax + q]) == 9 + q;
            if (reg && mine.size() > (size_t)gmax) reg = tb_of(mine[mine.size() - gmax - 1]) < 9;
            if (reg) nchain = mine.size() - gmax;
            for (size_t i = 0; i < nchain;) {
                size_t j = i + 1;
                while (j < nchain && tb_of(mine[j]) == tb_of(mine[j - 1]) + 1) ++j;
                const int len = (int)(j - i), ng = (len + gmax - 1) / gmax;
                size_t at = i;
                for (int q = 0; q < ng; ++q) {
                    const int sz = len / ng + (q < len % ng ? 1 : 0);
                    groups.push_back({at, sz});
                    at += sz;
                }
                i = j;
            }
            if (reg) groups.push_back({nchain, gmax});
            Launch l;
            l.d = dk;
            std::vector<int> chunk;
            auto emit = [&]() {
                if (chunk.empty()) return;
                l.d.nseg = (int)chunk.size();
                fwd.push_back(l);
                run_chunks.push_back(chunk);
                last = (int)fwd.size() - 1;
                l = Launch{};
                l.d = dk;
                chunk.clear();
            };
            for (const auto& gr : groups) {
                if (l.d.ngroups == kGroupMax || (int)chunk.size() + gr.second > kPassSegMax) emit();
                if (reg && gr.first == nchain) l.d.greg = l.d.ngroups;
                l.d.gt0[l.d.ngroups] = (uint32_t)tb_of(mine[gr.first]);
                l.d.gsize[l.d.ngroups] = (uint32_t)gr.second;
                l.d.gfirst[l.d.ngroups] = (uint32_t)chunk.size();
                ++l.d.ngroups;
                for (int q = 0; q < gr.second; ++q) chunk.push_back(mine[gr.first + q]);
            }
            emit();
        }
        if (st.cz >= 0) {
            if (last < 0) {  // a CZ set with no runs in its stage
                Launch l;
                l.d = pass_desc(g, c.n, 0);
                fwd.push_back(l);
                run_chunks.push_back({});
                last = (int)fwd.size() - 1;
            }
            set_cz(fwd[last].d, c.segs[st.cz].edges);
        }
    }
    std::vector<int> order(fwd.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = reverse ? (int)(order.size() - 1 - i) : (int)i;
    for (int i : order) {
        Launch l = fwd[i];
        if (l.tile) {
            l.d.first = (int)segv.size();
            l.d.nseg = (int)run_chunks[i].size();
            // groups are tile-bit intervals; one overlapping an earlier group
            // shares a qubit with it
            l.d.gundo = 0;
            for (int g2 = 0; g2 < l.d.ngroups; ++g2)
                for (int h = 0; h < g2; ++h)
                    if (l.d.gt0[h] < l.d.gt0[g2] + l.d.gsize[g2] && l.d.gt0[g2] < l.d.gt0[h] + l.d.gsize[h])
                        l.d.gundo |= 1u << g2;
            const int k = l.d.m == 0 ? 0 : pass_of(g, c.n, l.d.g0);
            for (int ri : run_chunks[i]) fill(c.segs[ri], l.d, k);
            split_pass(l.d, segv.data() + l.d.first, reverse, tabs);
        } else {
            l.d.greal = 0;
            l.d.tpre = l.d.tpost = -1;
        }
        launches.push_back(l);
    }
}

// Observable over tile passes: each flip-mask group goes to the first pass
// whose tile bits hold its flip mask; groups no pass holds (or with more terms
// than one launch takes) stay in the global-gather table `far`.  Inputs are
// the grouped table of iq_circuit_set_observable.
// Plans the Pauli sum (terms grouped by flip mask xm[g], group g owning terms
// [start[g], start[g+1])) over the tile passes of an n-qubit state of `dim`
// amplitudes; `p` must be empty (released).
bool plan_observable_tiles(ObsPlan& p, int n, int prec, uint64_t dim, const std::vector<uint64_t>& xm,
                           const std::vector<int>& start, const std::vector<uint64_t>& zy, const std::vector<int>& ny,
                           const std::vector<double>& co) {
    const TileGeom g = tile_geom(prec);
    const int npass = high_passes(g, n) + 1;
    std::vector<PassDesc> pd(npass);
    for (int k = 0; k < npass; ++k) pd[k] = pass_desc(g, n, k);
    auto split = [&](const PassDesc& d, uint64_t m, uint32_t& tile, uint64_t& fixed) {
        tile = 0;
        fixed = 0;
        for (int q = 0; q < n; ++q)
            if ((m >> q) & 1) {
                const int lb = local_bit(d, q);
                if (lb >= 0) tile |= 1u << lb;
                else fixed |= 1ULL << q;
            }
    };
    std::vector<std::vector<int>> by_pass(npass);
    std::vector<int> far;
    for (size_t gi = 0; gi < xm.size(); ++gi) {
        int at = -1;
        for (int k = 0; k < npass && at < 0; ++k) {
            uint32_t tl;
            uint64_t fx;
            split(pd[k], xm[gi], tl, fx);
            if (fx == 0) at = k;
        }
        if (at < 0 || start[gi + 1] - start[gi] > kObsTermsMax) far.push_back((int)gi);
        else by_pass[at].push_back((int)gi);
    }
    std::vector<ObsGroupT> G;
    std::vector<ObsTermT> Tm;
    std::vector<ObsChainT> Ch;
    p.opass.clear();
    for (int k = 0; k < npass; ++k) {
        ObsPass op{};
        op.d = pd[k];
        auto flush = [&]() {
            if (op.gcount > 0) p.opass.push_back(op);
            op.gfirst = (int)G.size();
            op.tfirst = (int)Tm.size();
            op.gcount = op.tcount = 0;
        };
        flush();
        for (int gi : by_pass[k]) {
            const int nt = start[gi + 1] - start[gi];
            if (op.tcount + nt > kObsTermsMax) flush();
            ObsGroupT gg{};
            uint64_t fx;
            split(pd[k], xm[gi], gg.xl, fx);
            gg.first = op.tcount;
            gg.count = nt;
            gg.nre = gg.nre0 = gg.nim0 = gg.nch = 0;
            gg.cfirst = (int)Ch.size();
            gg.ccount = 0;
            // the real-part run, then the imaginary-part run; in each, the
            // terms without sign bits on the tile first; in the real run the
            // chain-class terms next, class by class
            std::vector<ObsTermT> re0, rech, reot, im0, imot;
            std::vector<int> chain_of;  // class (relative) of each chain term, in re order
            for (int t = start[gi]; t < start[gi + 1]; ++t) {
                const int cls = ny[t] & 3;
                ObsTermT tt{};
                split(pd[k], zy[t], tt.zl, tt.zfix);
                tt.coeff = cls >= 2 ? -co[t] : co[t];
                if (cls & 1) {
                    (tt.zl ? imot : im0).push_back(tt);
                    continue;
                }
                if (!tt.zl) {
                    re0.push_back(tt);
                    continue;
                }
                // chain eligibility: one or two tile bits, none fixed by the tile
                const int nb = __builtin_popcount(tt.zl);
                int placed = -1;
                if (tt.zfix == 0 && nb <= 2) {
                    const uint32_t a = (uint32_t)__builtin_ctz(tt.zl);
                    const int sh = nb == 2 ? 31 - __builtin_clz(tt.zl) - (int)a : 0;
                    for (int c = gg.cfirst; c < (int)Ch.size() && placed < 0; ++c)
                        if (Ch[c].shift == sh && Ch[c].coeff == tt.coeff && !((Ch[c].mask >> a) & 1u)) placed = c;
                    if (placed < 0) {
                        ObsChainT nc{};
                        nc.shift = sh;
                        nc.coeff = tt.coeff;
                        Ch.push_back(nc);
                        placed = (int)Ch.size() - 1;
                    }
                    Ch[placed].mask |= 1u << a;
                    Ch[placed].count += 1;
                    rech.push_back(tt);
                    chain_of.push_back(placed);
                } else {
                    reot.push_back(tt);
                }
            }
            gg.ccount = (int)Ch.size() - gg.cfirst;
            for (const ObsTermT& t : re0) Tm.push_back(t);
            for (int c = gg.cfirst; c < (int)Ch.size(); ++c)  // chain terms class by class
                for (size_t i = 0; i < rech.size(); ++i)
                    if (chain_of[i] == c) Tm.push_back(rech[i]);
            for (const ObsTermT& t : reot) Tm.push_back(t);
            for (const ObsTermT& t : im0) Tm.push_back(t);
            for (const ObsTermT& t : imot) Tm.push_back(t);
            gg.nre0 = (int)re0.size();
            gg.nch = (int)rech.size();
            gg.nre = (int)(re0.size() + rech.size() + reot.size());
            gg.nim0 = (int)im0.size();
            G.push_back(gg);
            ++op.gcount;
            op.tcount += nt;
        }
        flush();
    }
    // far groups as their own flip-mask table
    std::vector<uint64_t> fxm, fzy;
    std::vector<int> fstart, fny;
    std::vector<double> fco;
    for (int gi : far) {
        fxm.push_back(xm[gi]);
        fstart.push_back((int)fzy.size());
        for (int t = start[gi]; t < start[gi + 1]; ++t) {
            fzy.push_back(zy[t]);
            fny.push_back(ny[t]);
            fco.push_back(co[t]);
        }
    }
    fstart.push_back((int)fzy.size());
    if (p.otab_dev) hipFree(p.otab_dev);
    if (p.far_dev) hipFree(p.far_dev);
    if (p.opart) hipFree(p.opart);
    p.otab_dev = p.far_dev = nullptr;
    p.opart = nullptr;
    p.far = ObsTable{};
    p.obs_tiled = false;
    const size_t gbytes = std::max<size_t>(16, G.size() * sizeof(ObsGroupT));
    const size_t tbytes = std::max<size_t>(16, Tm.size() * sizeof(ObsTermT));
    const size_t cbytes = std::max<size_t>(16, Ch.size() * sizeof(ObsChainT));
    const size_t toff = (gbytes + 255) & ~size_t(255), coff = (toff + tbytes + 255) & ~size_t(255);
    if (!hip_ok(hipMalloc(&p.otab_dev, coff + cbytes), "observable tiles")) return false;
    char* od = static_cast<char*>(p.otab_dev);
    if ((!G.empty() && !hip_ok(hipMemcpy(od, G.data(), G.size() * sizeof(ObsGroupT), hipMemcpyHostToDevice), "upload")) ||
        (!Tm.empty() && !hip_ok(hipMemcpy(od + toff, Tm.data(), Tm.size() * sizeof(ObsTermT), hipMemcpyHostToDevice), "upload")) ||
        (!Ch.empty() && !hip_ok(hipMemcpy(od + coff, Ch.data(), Ch.size() * sizeof(ObsChainT), hipMemcpyHostToDevice), "upload")))
        return false;
    p.ogroups = reinterpret_cast<const ObsGroupT*>(od);
    p.oterms = reinterpret_cast<const ObsTermT*>(od + toff);
    p.ochains = reinterpret_cast<const ObsChainT*>(od + coff);
    size_t parts = 0;
    for (const ObsPass& op : p.opass) parts += std::min<uint64_t>(op.d.ntiles, 2048);
    if (!far.empty()) {
        const size_t F = fxm.size(), FT = fzy.size();
        const size_t o1 = 0, o2 = o1 + F * 8, o3 = (o2 + (F + 1) * 4 + 7) & ~size_t(7), o4 = o3 + FT * 8,
                     o5 = (o4 + FT * 4 + 7) & ~size_t(7), bytes = o5 + FT * 8;
        std::vector<char> host(bytes, 0);
        std::memcpy(host.data() + o1, fxm.data(), F * 8);
        std::memcpy(host.data() + o2, fstart.data(), (F + 1) * 4);
        std::memcpy(host.data() + o3, fzy.data(), FT * 8);
        std::memcpy(host.data() + o4, fny.data(), FT * 4);
        std::memcpy(host.data() + o5, fco.data(), FT * 8);
        if (!hip_ok(hipMalloc(&p.far_dev, bytes), "observable table") ||
            !hip_ok(hipMemcpy(p.far_dev, host.data(), bytes, hipMemcpyHostToDevice), "observable upload"))
            return false;
        char* d = static_cast<char*>(p.far_dev);
        p.far.xmask = reinterpret_cast<const uint64_t*>(d + o1);
        p.far.start = reinterpret_cast<const int*>(d + o2);
        p.far.zy = reinterpret_cast<const uint64_t*>(d + o3);
        p.far.ny = reinterpret_cast<const int*>(d + o4);
        p.far.coeff = reinterpret_cast<const double*>(d + o5);
        p.far.ngroups = (int)F;
        parts += nblocks_for(dim);
    }
    // two per block: a statevector Pauli sum keeps Re and Im
    if (!hip_ok(hipMalloc((void**)&p.opart, 2 * std::max<size_t>(parts, 1) * sizeof(double)), "energy partials"))
        return false;
    p.opart_len = parts;
    p.obs_tiled = true;
    return true;
}

// Run descriptors, then (at a 16-byte boundary, `tab_off` bytes in) the split
// passes' phase tables as (re, im) pairs.
Circuit::DescSlot* upload_descs(Circuit& c, const std::vector<PassSeg>& segv, const std::vector<double>& tabs,
                                size_t& tab_off) {
    Circuit::DescSlot& d = c.dring[c.dnext];
    c.dnext = (c.dnext + 1) % 4;
    if (d.pending) hip_ok(hipEventSynchronize(d.ev), "descriptor wait");
    d.pending = false;
    tab_off = (segv.size() * sizeof(PassSeg) + 15) & ~size_t(15);
    const size_t bytes = std::max<size_t>(16, (tab_off + tabs.size() * sizeof(double) + 15) & ~size_t(15));
    if (bytes > d.cap) {
        if (d.host) hipHostFree(d.host);
        if (d.dev) hipFree(d.dev);
        d.host = d.host_dev = d.dev = nullptr;
        d.cap = 0;
        if (!hip_ok(hipHostMalloc(&d.host, bytes, hipHostMallocCoherent), "pinned alloc") ||
            !hip_ok(hipHostGetDevicePointer(&d.host_dev, d.host, 0), "pinned map") ||
            !hip_ok(hipMalloc(&d.dev, bytes), "alloc"))
            return nullptr;
        d.cap = bytes;
    }
    if (!d.ev && !hip_ok(hipEventCreateWithFlags(&d.ev, hipEventDisableTiming), "event")) return nullptr;
    std::memcpy(d.host, segv.data(), segv.size() * sizeof(PassSeg));
    if (!tabs.empty()) std::memcpy(static_cast<char*>(d.host) + tab_off, tabs.data(), tabs.size() * sizeof(double));
    if (!hip_ok(launch_upload(d.dev, d.host_dev, bytes, c.eng.stream), "descriptor upload")) return nullptr;
    return &d;
}

// A two-qubit gate of the tape, U (forward) or U^dagger (reverse sweep; CX is
// its own inverse), on each of the states a and b that is not null.
bool launch_two_qubit(Circuit& c, const Segment& s, bool dagger, void* a, void* b) {
    Engine& e = c.eng;
    const TapeGate& tg = c.gates[s.gate];
    double re[16], im[16];
    if (tg.kind == kCX) {
        const double cx[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0};
        std::memcpy(re, cx, sizeof(re));
        std::memset(im, 0, sizeof(im));
    } else if (!dagger) {
        std::memcpy(re, tg.mre, sizeof(re));
        std::memcpy(im, tg.mim, sizeof(im));
    } else {
        for (int r = 0; r < 4; ++r)
            for (int col = 0; col < 4; ++col) {
                re[r * 4 + col] = tg.mre[col * 4 + r];
                im[r * 4 + col] = -tg.mim[col * 4 + r];
            }
    }
    GateDesc g;
    g.K = 2;
    g.q[0] = (uint32_t)tg.q0;
    g.q[1] = (uint32_t)tg.q1;
    g.re = re;
    g.im = im;
    if (a && !hip_ok(launch_gate(a, c.n, c.prec, e.stream, g), "two-qubit gate")) return false;
    return !b || hip_ok(launch_gate(b, c.n, c.prec, e.stream, g), "two-qubit gate on lambda");
}

void release_stored(Circuit& c) {
    if (!c.stored) return;
    hipStreamSynchronize(c.eng.stream);
    hipFree(c.stored);
    c.stored = nullptr;
    c.stored_cap = 0;
    c.stored_valid = false;
}

// Room for the outputs of a tiled forward sweep of `nl` passes (the last one
// goes to the engine state): false when they exceed the circuit's budget or
// cannot be allocated, and the reverse sweep then takes psi back instead.
bool ensure_stored(Circuit& c, size_t nl) {
    const size_t need = (nl > 1 ? nl - 1 : 0) * (c.eng.dim * c.eng.elem);
    if (need > c.stored_budget) return false;
    if (need <= c.stored_cap) return true;
    release_stored(c);
    if (hipMalloc(&c.stored, need) != hipSuccess) {
        (void)hipGetLastError();
        c.stored = nullptr;
        return false;
    }
    c.stored_cap = need;
    return true;
}

// Where pass i of an nl-pass sweep writes: its stored slot, the last pass
// (and every pass when nothing is stored) the engine state.
void* pass_out(Circuit& c, bool stored, size_t i, size_t nl) {
    return stored && i + 1 < nl ? static_cast<char*>(c.stored) + i * (c.eng.dim * c.eng.elem) : c.eng.amp;
}

// Forward sweep; with `store`, every pass but the last writes its own stored
// slot (same traffic as in place), for the reverse sweep to read.
// A tape's forward plan, kept when no gate takes a parameter.
struct FixedPlan {
    std::vector<PassSeg> segv;
    std::vector<Launch> launches;
    std::vector<double> tabs;
};

template <typename R>
bool run_forward_tiled(Circuit& c, const double* params, bool store) {
    Engine& e = c.eng;
    c.stored_valid = false;
    if (!e.reset_basis(0)) return false;
    bool fixed = true;
    for (const auto& g : c.gates) fixed = fixed && g.p[0] < 0;
    if (fixed && !c.fixed_plan) {
        auto fp = std::make_shared<FixedPlan>();
        plan_tiles(c, params, false, fp->segv, fp->launches, fp->tabs);
        c.fixed_plan = fp;
    }
    std::vector<PassSeg> seg_local;
    std::vector<Launch> launch_local;
    std::vector<double> tab_local;
    if (!fixed) plan_tiles(c, params, false, seg_local, launch_local, tab_local);
    const FixedPlan* fp = fixed ? static_cast<const FixedPlan*>(c.fixed_plan.get()) : nullptr;
    const std::vector<PassSeg>& segv = fp ? fp->segv : seg_local;
    const std::vector<Launch>& launches = fp ? fp->launches : launch_local;
    const std::vector<double>& tabs = fp ? fp->tabs : tab_local;
    size_t tab_off = 0;
    Circuit::DescSlot* ds = upload_descs(c, segv, tabs, tab_off);
    if (!ds) return false;
    const PassSeg* dsegs = static_cast<const PassSeg*>(ds->dev);
    const double2* dtabs = reinterpret_cast<const double2*>(static_cast<const char*>(ds->dev) + tab_off);
    const size_t nl = launches.size(), bytes = e.dim * e.elem;
    const bool st = store && ensure_stored(c, nl);
    for (size_t i = 0; i < nl; ++i) {
        const Launch& l = launches[i];
        void* src = i == 0 ? e.amp : pass_out(c, st, i - 1, nl);
        void* dst = pass_out(c, st, i, nl);
        if (!l.tile) {
            TimedLaunch tl(c, kLaunchTwoQubit);
            if (src != dst && !hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, e.stream), "state copy"))
                return false;
            if (!launch_two_qubit(c, c.segs[l.seg2], false, dst, nullptr)) return false;
            continue;
        }
        const size_t lds = (((size_t)1 << l.d.ts_bits) + 2 * kPhaseTab) * sizeof(cv<R>);
        const unsigned grid = (unsigned)std::min<uint64_t>(l.d.ntiles, 2048);
        hipFuncSetAttribute((const void*)k_tile_fwd<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        TimedLaunch tl(c, kLaunchForward);
        hipLaunchKernelGGL(k_tile_fwd<R>, dim3(grid), dim3(512), lds, e.stream, (const cv<R>*)src, (cv<R>*)dst, l.d,
                           dsegs, dtabs);
        if (!hip_ok(hipGetLastError(), "tiled forward")) return false;
    }
    ds->pending = hip_ok(hipEventRecord(ds->ev, e.stream), "event record");
    c.stored_valid = ds->pending && st;
    return ds->pending;
}

template <typename R>
bool run_reverse_tiled(Circuit& c, const double* params) {
    Engine& e = c.eng;
    std::vector<PassSeg> segv;
    std::vector<Launch> launches;
    std::vector<double> tabs;
    plan_tiles(c, params, true, segv, launches, tabs);
    size_t tab_off = 0;
    Circuit::DescSlot* ds = upload_descs(c, segv, tabs, tab_off);
    if (!ds) return false;
    const PassSeg* dsegs = static_cast<const PassSeg*>(ds->dev);
    const double2* dtabs = reinterpret_cast<const double2*>(static_cast<const char*>(ds->dev) + tab_off);
    // with the forward sweep's stored outputs, psi is read, never taken back
    const bool st = c.stored_valid;
    const size_t nl = launches.size();
    for (size_t k = 0; k < nl; ++k) {
        const Launch& l = launches[k];
        const size_t i = nl - 1 - k;  // the pass's place in the forward sweep
        if (!l.tile) {
            TimedLaunch tl(c, kLaunchTwoQubit);
            if (!launch_two_qubit(c, c.segs[l.seg2], true, st ? nullptr : e.amp, c.lam)) return false;
            continue;
        }
        const size_t wred_bytes = (size_t)8 * kPassSegMax * kSubMax * sizeof(double);
        if (st && l.d.gundo == 0 && c.rev_lam_only) {
            // psi read from its stored slot, lambda alone in LDS
            const size_t lds = ((size_t)1 << l.d.ts_bits) * sizeof(cv<R>) + wred_bytes + 2 * kPhaseTab * sizeof(cv<R>);
            hipFuncSetAttribute((const void*)k_tile_rev_lam<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            TimedLaunch tl(c, kLaunchReverse);
            hipLaunchKernelGGL(k_tile_rev_lam<R>, dim3(c.nblk), dim3(512), lds, e.stream,
                               (const cv<R>*)pass_out(c, st, i, nl), (cv<R>*)c.lam, l.d, dsegs, dtabs, c.gpart,
                               (int)c.nblk);
            if (!hip_ok(hipGetLastError(), "tiled reverse")) return false;
            continue;
        }
        const size_t lds = 2 * ((size_t)1 << l.d.ts_bits) * sizeof(cv<R>) + wred_bytes + 2 * kPhaseTab * sizeof(cv<R>);
        auto kern = st ? k_tile_rev<R, true> : k_tile_rev<R, false>;
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        TimedLaunch tl(c, kLaunchReverse);
        hipLaunchKernelGGL(kern, dim3(c.nblk), dim3(512), lds, e.stream, (cv<R>*)pass_out(c, st, i, nl),
                           (cv<R>*)c.lam, l.d, dsegs, dtabs, c.gpart, (int)c.nblk);
        if (!hip_ok(hipGetLastError(), "tiled reverse")) return false;
    }
    c.stored_valid = false;
    if (c.nslots > 0) {
        TimedLaunch tl(c, kLaunchFinalize);
        hipLaunchKernelGGL(k_finalize_slots, dim3(c.nslots * kSubMax), dim3(256), 0, e.stream, c.gpart, (int)c.nblk,
                           c.nslots, c.gout);
        if (!hip_ok(hipGetLastError(), "gradient finalize")) return false;
    }
    ds->pending = hip_ok(hipEventRecord(ds->ev, e.stream), "event record");
    return ds->pending;
}

template <typename R>
bool run_forward(Circuit& c, const double* params, bool store = false) {
    return use_tiles(c) ? run_forward_tiled<R>(c, params, store) : run_forward_segments<R>(c, params);
}

template <typename R>
bool run_reverse(Circuit& c, const double* params) {
    return use_tiles(c) ? run_reverse_tiled<R>(c, params) : run_reverse_segments<R>(c, params);
}

Circuit* circ_of(void* p) { return static_cast<Circuit*>(p); }

bool prepare(Circuit& c, const double* params, size_t nparams) {
    if (c.terms.empty()) {
        set_error(kErrInvalidArgument, "no observable set");
        return false;
    }
    if (!params && nparams) {
        set_error(kErrInvalidArgument, "null parameter array");
        return false;
    }
    if (!check_params(c, nparams)) return false;
    if (!c.compiled) compile(c);
    return true;
}

}  // namespace

// Sum_t coeff_t <psi|P_t|psi> over a statevector handle's state through
// the tile passes, with the reference's per-index skip
// (core/src/statevector.cpp:258-291 term by term).  One pass per flip mask
// (up to eight strings each) costs a read of the state per mask; the tile
// passes read it once per pass (three at 30 qubits complex128) and take
// every mask on the pass's tile bits.  Sets *used = false, touching nothing,
// when the tile passes would not save a read (small states, few masks).
bool pauli_sum_tiled(Engine& e, size_t nterms, const PauliTerm* terms, const double* coeffs, double* out, bool* used) {
    // out[0], out[1]: Re and Im
    *used = false;
    if (e.n < 20 || e.n > 34 || nterms == 0) return true;
    std::vector<size_t> order(nterms);
    for (size_t i = 0; i < nterms; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return terms[a].xmask < terms[b].xmask; });
    std::vector<uint64_t> xm, zy;
    std::vector<int> start, ny;
    std::vector<double> co;
    for (size_t k = 0; k < nterms; ++k) {
        const PauliTerm& t = terms[order[k]];
        if (xm.empty() || xm.back() != t.xmask) {
            xm.push_back(t.xmask);
            start.push_back((int)zy.size());
        }
        zy.push_back(t.zymask);
        ny.push_back(t.ny);
        co.push_back(coeffs ? coeffs[order[k]] : 1.0);
    }
    start.push_back((int)zy.size());
    size_t mask_passes = 0;  // of the per-mask reduction, eight strings per pass
    for (size_t gi = 0; gi + 1 < start.size(); ++gi) mask_passes += (size_t)(start[gi + 1] - start[gi] + 7) / 8;
    const TileGeom g = tile_geom(e.prec);
    const int npass = high_passes(g, e.n) + 1;
    if (mask_passes <= (size_t)npass) return true;
    ObsPlan plan;
    if (!plan_observable_tiles(plan, e.n, e.prec, e.dim, xm, start, zy, ny, co)) return false;
    const bool ok = e.prec == kC128
                        ? launch_observable_plan<double>(plan, e.amp, nullptr, e.dim, e.stream, false, true, e.red_out)
                        : launch_observable_plan<float>(plan, e.amp, nullptr, e.dim, e.stream, false, true, e.red_out);
    // the plan's tables are freed after the stream has finished with them
    if (!ok || !e.fetch(2, out)) {
        e.sync();
        return false;
    }
    *used = true;
    return true;
}

}  // namespace iqg

using namespace iqg;

extern "C" {

void* iq_circuit_create(size_t num_qubits, int precision) {
    if (num_qubits < 1 || num_qubits > 34 || (precision != kC128 && precision != kC64)) {
        set_error(kErrInvalidArgument, "unsupported circuit shape (n=%zu, precision=%d)", num_qubits, precision);
        return nullptr;
    }
    Circuit* c = new (std::nothrow) Circuit();
    if (!c) return nullptr;
    c->n = (int)num_qubits;
    c->prec = precision;
    if (!c->eng.init((int)num_qubits, precision) || hipMalloc(&c->lam, c->eng.dim * c->eng.elem) != hipSuccess) {
        if (!c->lam) set_error(kErrOutOfMemory, "cannot allocate the adjoint state");
        delete c;
        return nullptr;
    }
    c->nblk = nblocks_for(c->eng.dim >> 1);
    return c;
}

void iq_circuit_free(void* c) {
    if (!c) return;
    DeviceGuard g(circ_of(c)->eng.dev);
    delete circ_of(c);
}

// kind: 0 RX, 1 RY, 2 RZ (param_index[0]); 3 U3 (param_index[0..2] = theta, phi, lambda);
// 4 CZ, 5 CX (q0 control); 6 fixed 2x2 matrix; 7 fixed 4x4 matrix (MSB-first on q0, q1).
int iq_circuit_add_gate(void* cp, int kind, size_t q0, size_t q1, const int64_t* param_index, const double* m_re,
                        const double* m_im) {
    if (!cp) return -1;
    Circuit& c = *circ_of(cp);
    TapeGate g;
    g.kind = kind;
    g.q0 = (int)std::min<size_t>(q0, 1 << 20);
    g.q1 = (int)std::min<size_t>(q1, 1 << 20);
    const bool two = (kind == kCZ || kind == kCX || kind == kMat2);
    if (kind < 0 || kind > kMat2 || g.q0 >= c.n || (two && (g.q1 >= c.n || g.q1 == g.q0))) {
        set_error(kErrInvalidArgument, "invalid gate (kind %d, qubits %zu, %zu)", kind, q0, q1);
        return -1;
    }
    const int np = (kind == kU3) ? 3 : (kind <= kRZ ? 1 : 0);
    for (int i = 0; i < np; ++i) {
        if (!param_index || param_index[i] < 0) {
            set_error(kErrInvalidArgument, "parametric gate needs non-negative parameter indices");
            return -1;
        }
        g.p[i] = param_index[i];
    }
    if (kind == kMat1 || kind == kMat2) {
        if (!m_re || !m_im) {
            set_error(kErrInvalidArgument, "fixed-matrix gate needs a matrix");
            return -1;
        }
        const int d = kind == kMat1 ? 4 : 16;
        for (int i = 0; i < d; ++i) {
            g.mre[i] = m_re[i];
            g.mim[i] = m_im[i];
        }
    }
    c.gates.push_back(g);
    c.compiled = false;
    c.fixed_plan.reset();
    return (int)c.gates.size() - 1;
}

size_t iq_circuit_num_gates(void* c) { return c ? circ_of(c)->gates.size() : 0; }

// Number of HBM passes of the compiled tape's forward sweep: tile stages
// (and standalone two-qubit gates) when tiled, else segments.  The reverse
// sweep runs the same passes backwards, each touching both states.  The plan
// does not depend on parameter values, so it is counted with zeros.
size_t iq_circuit_num_segments(void* cp) {
    if (!cp) return 0;
    Circuit& c = *circ_of(cp);
    if (!c.compiled) compile(c);
    if (!use_tiles(c)) return c.segs.size();
    int64_t np = 0;
    for (const TapeGate& g : c.gates)
        for (int64_t p : g.p) np = std::max(np, p + 1);
    const std::vector<double> zeros((size_t)np, 0.0);
    std::vector<PassSeg> segv;
    std::vector<Launch> launches;
    std::vector<double> tabs;
    plan_tiles(c, zeros.data(), false, segv, launches, tabs);
    return launches.size();
}

int iq_circuit_set_observable(void* cp, size_t nterms, const double* coeffs, const char* ops, const size_t* qubits,
                              const size_t* lens) {
    if (!cp || !ops || !qubits || !lens || nterms == 0) return -1;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    std::vector<PauliTerm> terms(nterms);
    size_t off = 0;
    for (size_t t = 0; t < nterms; ++t) {
        if (!parse_pauli(ops + off, qubits + off, lens[t], c.n, terms[t])) return -1;
        off += lens[t];
    }
    // group by flip mask
    std::vector<size_t> order(nterms);
    for (size_t i = 0; i < nterms; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return terms[a].xmask < terms[b].xmask; });
    std::vector<uint64_t> xm, zy;
    std::vector<int> start, ny;
    std::vector<double> co;
    for (size_t k = 0; k < nterms; ++k) {
        const PauliTerm& t = terms[order[k]];
        if (xm.empty() || xm.back() != t.xmask) {
            xm.push_back(t.xmask);
            start.push_back((int)zy.size());
        }
        zy.push_back(t.zymask);
        ny.push_back(t.ny);
        co.push_back(coeffs ? coeffs[order[k]] : 1.0);
    }
    start.push_back((int)zy.size());
    const size_t G = xm.size(), T = zy.size();
    const size_t bytes = G * 8 + (G + 1) * 4 + T * 8 + T * 4 + T * 8 + 64;
    if (c.obs_dev) hipFree(c.obs_dev);
    c.obs_dev = nullptr;
    if (!hip_ok(hipMalloc(&c.obs_dev, bytes), "observable table")) return -1;
    std::vector<char> host(bytes, 0);
    char* base = host.data();
    size_t o = 0;
    auto put = [&](const void* src, size_t n) {
        std::memcpy(base + o, src, n);
        const size_t at = o;
        o += (n + 7) & ~size_t(7);
        return at;
    };
    const size_t oxm = put(xm.data(), G * 8), ost = put(start.data(), (G + 1) * 4), ozy = put(zy.data(), T * 8),
                 ony = put(ny.data(), T * 4), oco = put(co.data(), T * 8);
    if (!hip_ok(hipMemcpy(c.obs_dev, host.data(), bytes, hipMemcpyHostToDevice), "observable upload")) return -1;
    char* d = static_cast<char*>(c.obs_dev);
    c.obs.xmask = reinterpret_cast<const uint64_t*>(d + oxm);
    c.obs.start = reinterpret_cast<const int*>(d + ost);
    c.obs.zy = reinterpret_cast<const uint64_t*>(d + ozy);
    c.obs.ny = reinterpret_cast<const int*>(d + ony);
    c.obs.coeff = reinterpret_cast<const double*>(d + oco);
    c.obs.ngroups = (int)G;
    c.terms = terms;
    c.oplan.release();
    if (use_tiles(c) && !plan_observable_tiles(c.oplan, c.n, c.prec, c.eng.dim, xm, start, zy, ny, co)) return -1;
    return 0;
}

double iq_circuit_expectation(void* cp, const double* params, size_t nparams) {
    if (!cp) return 0.0;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!prepare(c, params, nparams)) return 0.0;
    c.tused = 0;
    const bool ok = c.prec == kC128 ? (run_forward<double>(c, params) && apply_observable<double>(c, false, c.eng.red_out))
                                    : (run_forward<float>(c, params) && apply_observable<float>(c, false, c.eng.red_out));
    double e = 0.0;
    if (ok) c.eng.fetch(1, &e);
    return e;
}

// The forward sweep alone: the circuit's state becomes U(params)|0..0>, with
// no observable.  Asynchronous on the circuit's stream.
int iq_circuit_run(void* cp, const double* params, size_t nparams) {
    if (!cp || (!params && nparams)) {
        set_error(kErrInvalidArgument, "null circuit or parameter array");
        return -1;
    }
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!check_params(c, nparams)) return -1;
    if (!c.compiled) compile(c);
    c.tused = 0;
    const bool ok = c.prec == kC128 ? run_forward<double>(c, params) : run_forward<float>(c, params);
    return ok ? 0 : -1;
}

// Copies the circuit's state into statevector handle sv (same qubit count,
// precision and device), after the circuit's work; ordered on sv's stream.
int iq_circuit_copy_state(void* cp, void* svp) {
    if (!cp || !svp) {
        set_error(kErrInvalidArgument, "null circuit or statevector");
        return -1;
    }
    Circuit& c = *circ_of(cp);
    Engine& s = *static_cast<Engine*>(svp);
    if (s.n != c.n || s.prec != c.prec || s.dev != c.eng.dev) {
        set_error(kErrInvalidArgument, "statevector shape (n=%d, precision %d) differs from the circuit's (n=%d, %d)",
                  s.n, s.prec, c.n, c.prec);
        return -1;
    }
    DeviceGuard dg(c.eng.dev);
    if (!c.eng.sync()) return -1;
    if (!hip_ok(hipMemcpyAsync(s.amp, c.eng.amp, c.eng.dim * c.eng.elem, hipMemcpyDeviceToDevice, s.stream),
                "state copy"))
        return -1;
    return 0;
}

int iq_circuit_set_lambda_tiles(void* cp, int on) {
    if (!cp) return -1;
    circ_of(cp)->rev_lam_only = on != 0;
    return 0;
}

int iq_circuit_set_timing(void* cp, int on) {
    if (!cp) return -1;
    Circuit& c = *circ_of(cp);
    c.timing = on != 0;
    c.tused = 0;
    return 0;
}

size_t iq_circuit_launch_times(void* cp, int* kinds, double* ms, size_t cap) {
    if (!cp) return 0;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!c.eng.sync()) return 0;
    const size_t m = std::min(cap, c.tused);
    for (size_t i = 0; i < m; ++i) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, c.tev[2 * i], c.tev[2 * i + 1]) != hipSuccess) t = -1.f;
        if (kinds) kinds[i] = c.tkind[i];
        if (ms) ms[i] = t;
    }
    return c.tused;
}

// Adjoint gradient: grad_out[j] = d<H>/d params[j] (summed over every gate using j).
int iq_circuit_gradient(void* cp, const double* params, size_t nparams, double* energy_out, double* grad_out) {
    if (!cp || (!params && nparams) || !grad_out) return -1;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!prepare(c, params, nparams) || !ensure_grad_buffers(c)) return -1;
    c.tused = 0;
    const bool ok = c.prec == kC128
                        ? (run_forward<double>(c, params, true) && apply_observable<double>(c, true, c.eng.red_out) &&
                           run_reverse<double>(c, params))
                        : (run_forward<float>(c, params, true) && apply_observable<float>(c, true, c.eng.red_out) &&
                           run_reverse<float>(c, params));
    if (!ok) return -1;
    double e = 0.0;
    if (c.nslots > 0 &&
        !hip_ok(hipMemcpyAsync(c.hgout, c.gout, (size_t)c.nslots * kSubMax * 8, hipMemcpyDeviceToHost, c.eng.stream), "D2H"))
        return -1;
    if (!c.eng.fetch(1, &e)) return -1;
    for (size_t j = 0; j < nparams; ++j) grad_out[j] = 0.0;
    const bool tiled = use_tiles(c);
    for (const auto& s : c.segs) {
        if (s.kind != 0) continue;
        const double* slot = c.hgout + (size_t)s.slot * kSubMax;
        if (!tiled) {  // per sub-step gradients
            for (size_t k = 0; k < s.subs.size(); ++k) {
                const int64_t p = sub_param(c.gates[s.subs[k].gate], s.subs[k].which);
                if (p >= 0) grad_out[p] += slot[k];
            }
            continue;
        }
        // the run's overlaps T_X, T_Y, T_Z: sub-step k's gradient is
        // Im<lambda| V G V^dagger psi> with V the run's later sub-steps, and
        // V G V^dagger = wx X + wy Y + wz Z
        M2 V = {{1.0, 0.0, 0.0, 1.0}};
        for (int k = (int)s.subs.size() - 1; k >= 0; --k) {
            const TapeGate& tg = c.gates[s.subs[k].gate];
            const int64_t p = sub_param(tg, s.subs[k].which);
            if (p >= 0) {
                const M2 h = mul(mul(V, pauli(sub_gen(tg, s.subs[k].which))), dag(V));
                const double wx = h.m[1].real(), wy = -h.m[1].imag(), wz = 0.5 * (h.m[0].real() - h.m[3].real());
                grad_out[p] += wx * slot[0] + wy * slot[1] + wz * slot[2];
            }
            V = mul(V, sub_matrix(tg, s.subs[k].which, params));
        }
    }
    if (energy_out) *energy_out = e;
    return 0;
}

// Parameter-shift gradient on the same engine, for the parameters listed in
// `which` (all when which == NULL): (E(theta + pi/2) - E(theta - pi/2)) / 2,
// exact when each listed parameter drives one Pauli rotation.
int iq_circuit_parameter_shift(void* cp, const double* params, size_t nparams, const int64_t* which, size_t nwhich,
                               double* grad_out) {
    if (!cp || !params || !grad_out) return -1;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    if (!prepare(c, params, nparams)) return -1;
    const size_t count = which ? nwhich : nparams;
    if (c.evals_cap < 2 * count) {
        if (c.evals) hipFree(c.evals);
        c.evals = nullptr;
        c.evals_cap = 0;
        if (!hip_ok(hipMalloc(&c.evals, std::max<size_t>(2 * count, 1) * 8), "shift energies")) return -1;
        c.evals_cap = 2 * count;
    }
    std::vector<double> shifted(params, params + nparams);
    const double h = M_PI / 2;
    for (size_t k = 0; k < count; ++k) {
        const int64_t j = which ? which[k] : (int64_t)k;
        if (j < 0 || (size_t)j >= nparams) {
            set_error(kErrInvalidArgument, "parameter index out of range");
            return -1;
        }
        for (int sgn = 0; sgn < 2; ++sgn) {
            shifted[j] = params[j] + (sgn == 0 ? h : -h);
            const bool ok = c.prec == kC128
                                ? (run_forward<double>(c, shifted.data()) && apply_observable<double>(c, false, c.evals + 2 * k + sgn))
                                : (run_forward<float>(c, shifted.data()) && apply_observable<float>(c, false, c.evals + 2 * k + sgn));
            if (!ok) return -1;
        }
        shifted[j] = params[j];
    }
    std::vector<double> ev(2 * count);
    if (count && (!hip_ok(hipMemcpyAsync(ev.data(), c.evals, 2 * count * 8, hipMemcpyDeviceToHost, c.eng.stream), "D2H") ||
                  !c.eng.sync()))
        return -1;
    for (size_t k = 0; k < count; ++k) grad_out[k] = 0.5 * (ev[2 * k] - ev[2 * k + 1]);
    return 0;
}

// The engine state: after an expectation or parameter shift, the circuit's
// output state.  After a gradient, the output state when the forward sweep's
// pass outputs were stored (the reverse sweep only reads them), otherwise psi
// taken back to |0..0> up to rounding.
void iq_circuit_get_amplitudes(void* cp, double* re, double* im) {
    if (!cp || !re || !im) return;
    DeviceGuard dg(circ_of(cp)->eng.dev);
    circ_of(cp)->eng.export_split(re, im);
}

void* iq_circuit_stream(void* cp) { return cp ? (void*)circ_of(cp)->eng.stream : nullptr; }

int iq_circuit_set_stored_budget(void* cp, size_t bytes) {
    if (!cp) return -1;
    Circuit& c = *circ_of(cp);
    DeviceGuard dg(c.eng.dev);
    c.stored_budget = bytes;
    if (c.stored_cap > bytes) release_stored(c);
    return 0;
}

}  // extern "C"
