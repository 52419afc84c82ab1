// Index layout of the gate kernels, shared verbatim by the device code and a
// host-only (g++) checker so the address arithmetic is verified on the CPU.
//
// direct: group g (a compact index over the non-target qubits) starts at
//   insert_zeros(g, sorted targets); member s (MSB-first local index) is at
//   base + off[s].
// tiled:  a tile is 2^(K+6) amplitudes spanning the K targets plus the six
//   lowest non-target qubits (these always include qubits 0..5, so tile
//   element j = lane + 64*m sits at tile_base + lane + off[m]).  LDS position
//   of tile element j is j ^ swz(j): the bits of the lowest non-target tile
//   positions above 3 are XORed onto the target positions below 4, so both the
//   coalesced pass (16 consecutive j) and the gather pass (16 consecutive
//   groups) touch 16 distinct 16-element bank slots.
#pragma once

#include <cstdint>
#include <vector>
#include <algorithm>

#if defined(__HIPCC__) || defined(__HIP__)
#define IQ_HD __host__ __device__ __forceinline__
#else
#define IQ_HD inline
#endif

namespace iqg {

template <int K>
struct DirectArgs {
    uint32_t pos[K];         // targets ascending (zero-bit insertion)
    uint64_t off[1 << K];    // global offset of local index s
};

template <int K>
struct TiledArgs {
    uint64_t off[1 << K];  // global offset of tile element 64*m
    uint32_t pm[1 << K];   // LDS position contribution of tile element 64*m
    uint32_t ts[1 << K];   // tile-index target part of local index s
    uint64_t rest;         // global qubits outside the tile
    uint64_t ntiles;
    uint32_t ntmask;       // tile-index positions of the six non-target tile bits
    uint32_t swz[4];       // swizzle pairs: source position | destination position << 8
    uint32_t nswz;
};

template <int K>
IQ_HD uint64_t insert_zero_bits(uint64_t g, const uint32_t (&pos)[K]) {
    for (int m = 0; m < K; ++m) {
        const uint32_t p = pos[m];
        const uint64_t lo = g & ((1ULL << p) - 1);
        g = ((g >> p) << (p + 1)) | lo;
    }
    return g;
}

IQ_HD uint64_t deposit_bits(uint64_t v, uint64_t mask) {
    uint64_t out = 0;
    for (uint64_t bit = 1; mask; bit <<= 1) {
        const uint64_t low = mask & (~mask + 1);
        if (v & bit) out |= low;
        mask &= mask - 1;
    }
    return out;
}

IQ_HD uint32_t swizzle_bits(uint32_t x, const uint32_t (&swz)[4], uint32_t nswz) {
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i)
        if ((uint32_t)i < nswz) r |= ((x >> (swz[i] & 0xffu)) & 1u) << (swz[i] >> 8);
    return r;
}

// LDS positions of a lane: pl for the coalesced pass (XOR with pm[m]) and pg
// for its own group (XOR with ts[s]).
template <int K>
IQ_HD void tile_lane_positions(uint32_t lane, const TiledArgs<K>& a, uint32_t& pl, uint32_t& pg) {
    pl = lane ^ swizzle_bits(lane, a.swz, a.nswz);
    uint32_t g = 0, mask = a.ntmask;
    for (uint32_t b = 0; b < 6; ++b) {
        const uint32_t low = mask & (~mask + 1u);
        if ((lane >> b) & 1u) g |= low;
        mask &= mask - 1u;
    }
    pg = g ^ swizzle_bits(g, a.swz, a.nswz);
}

template <int K>
inline void fill_direct_args(DirectArgs<K>& a, const uint32_t* q) {
    uint32_t s[8];
    for (int m = 0; m < K; ++m) s[m] = q[m];
    std::sort(s, s + K);
    for (int m = 0; m < K; ++m) a.pos[m] = s[m];
    for (int l = 0; l < (1 << K); ++l) {
        uint64_t o = 0;
        for (int m = 0; m < K; ++m)
            if ((l >> (K - 1 - m)) & 1) o |= 1ULL << q[m];
        a.off[l] = o;
    }
}

// MFMA kernel for 5-qubit blocks (16x16x4 matrix-core tiles).  One wave holds
// 16 groups (lane & 15 = group within the batch); the four lanes of a group
// (lane >> 4 = q) hold 8 of its 32 amplitudes each.  Column kappa = 4*kb + q of
// the matrix tiles reads the amplitude at offset xq[q] + xk[kb]: q spans the two
// lowest target qubits, kb the other three, so global accesses stay in runs.
// Output value jj of lane (q, n) in row tile mt is the amplitude at offset
// yq[q] + yk[4*mt + jj] (q -> two lowest targets, jj -> next two, mt -> top).
// Its tile row is 16*mt + 4*q + jj for the f32 16x16x4 instruction and
// 16*mt + q + 4*jj for the f64 one (measured lane layouts of the two).
struct MfmaArgs {
    uint32_t pos[5];      // targets ascending (zero-bit insertion)
    uint64_t xq[4], xk[8];
    uint64_t yq[4], yk[8];
    uint64_t nbatches;    // 2^(n-5) / 16
    uint32_t rot, bits;   // visit order of super-batches: rotate_index(i, rot, bits)
    uint32_t xm;          // XCD-consistent split (one super-batch per wave, 2^xm of them), 0: off
};

// Left rotation of i inside its low `bits` bits: a bijection on [0, 2^bits).
// With rot = r, waves that run at the same time take super-batches 2^r apart.
IQ_HD uint64_t rotate_index(uint64_t i, uint32_t rot, uint32_t bits) {
    if (rot == 0 || rot >= bits) return i;
    const uint64_t mask = (1ULL << bits) - 1;
    return ((i << rot) | (i >> (bits - rot))) & mask;
}

// Local (MSB-first) index of tile column kappa / tile row rho, used to permute
// the matrix into MFMA fragment order on the host.
IQ_HD uint32_t mfma_row(uint32_t q, uint32_t jj, uint32_t mt, bool f64) {
    return 16u * mt + (f64 ? q + 4u * jj : 4u * q + jj);
}

inline void fill_mfma_args(MfmaArgs& a, int n, const uint32_t* q, bool f64, uint32_t* s_in, uint32_t* s_out) {
    uint32_t t[5];
    int lbit[5];  // local bit of target t[b]
    for (int m = 0; m < 5; ++m) t[m] = q[m];
    std::sort(t, t + 5);
    for (int b = 0; b < 5; ++b)
        for (int m = 0; m < 5; ++m)
            if (q[m] == t[b]) lbit[b] = 4 - m;
    for (int m = 0; m < 5; ++m) a.pos[m] = t[m];
    // kappa bit b <-> target t[b]
    for (int qq = 0; qq < 4; ++qq) a.xq[qq] = ((qq & 1) ? 1ULL << t[0] : 0) | ((qq & 2) ? 1ULL << t[1] : 0);
    for (int kb = 0; kb < 8; ++kb)
        a.xk[kb] = ((kb & 1) ? 1ULL << t[2] : 0) | ((kb & 2) ? 1ULL << t[3] : 0) | ((kb & 4) ? 1ULL << t[4] : 0);
    for (uint32_t kap = 0; kap < 32; ++kap) {
        uint32_t s = 0;
        for (int b = 0; b < 5; ++b)
            if ((kap >> b) & 1u) s |= 1u << lbit[b];
        s_in[kap] = s;
    }
    // f32 rows: bits 0,1 (jj) <-> t[2], t[3]; 2,3 (q) <-> t[0], t[1]; 4 (mt) <-> t[4]
    // f64 rows: bits 0,1 (q) <-> t[0], t[1]; 2,3 (jj) <-> t[2], t[3]; 4 (mt) <-> t[4]
    const int rb32[5] = {2, 3, 0, 1, 4}, rb64[5] = {0, 1, 2, 3, 4};
    const int* rb_of_target = f64 ? rb64 : rb32;
    for (int qq = 0; qq < 4; ++qq) a.yq[qq] = ((qq & 1) ? 1ULL << t[0] : 0) | ((qq & 2) ? 1ULL << t[1] : 0);
    for (int r = 0; r < 8; ++r) {  // r = 4*mt + jj
        const uint32_t jj = r & 3, mt = r >> 2;
        a.yk[r] = ((jj & 1) ? 1ULL << t[2] : 0) | ((jj & 2) ? 1ULL << t[3] : 0) | (mt ? 1ULL << t[4] : 0);
    }
    for (uint32_t rho = 0; rho < 32; ++rho) {
        uint32_t s = 0;
        for (int b = 0; b < 5; ++b)
            if ((rho >> rb_of_target[b]) & 1u) s |= 1u << lbit[b];
        s_out[rho] = s;
    }
    a.nbatches = (n >= 9) ? (1ULL << (n - 5)) / 16 : 0;
    a.rot = 0;
    a.bits = 0;
    a.xm = 0;
}

// Requires n >= K + 6 and distinct targets < n.
template <int K>
inline void fill_tiled_args(TiledArgs<K>& a, int n, const uint32_t* q) {
    uint64_t tmask = 0;
    for (int m = 0; m < K; ++m) tmask |= 1ULL << q[m];
    uint64_t bmask = tmask;
    int added = 0;
    for (int b = 0; b < n && added < 6; ++b)
        if (!((tmask >> b) & 1ULL)) {
            bmask |= 1ULL << b;
            ++added;
        }
    std::vector<int> bsorted;
    int tpos[64] = {0};
    for (int b = 0; b < n; ++b)
        if ((bmask >> b) & 1ULL) {
            tpos[b] = (int)bsorted.size();
            bsorted.push_back(b);
        }
    const int nb = (int)bsorted.size();
    auto to_global = [&](uint32_t j) {
        uint64_t o = 0;
        for (int p = 0; p < nb; ++p)
            if ((j >> p) & 1u) o |= 1ULL << bsorted[p];
        return o;
    };
    a.ntmask = 0;
    std::vector<int> ntpos;
    for (int p = 0; p < nb; ++p)
        if (!((tmask >> bsorted[p]) & 1ULL)) {
            a.ntmask |= 1u << p;
            ntpos.push_back(p);
        }
    std::vector<int> lhi, mpos;
    for (int i = 0; i < 4; ++i)
        if (ntpos[i] >= 4) lhi.push_back(ntpos[i]);
    for (int p = 0; p < 4; ++p)
        if ((tmask >> bsorted[p]) & 1ULL) mpos.push_back(p);
    a.nswz = (uint32_t)lhi.size();
    for (int i = 0; i < 4; ++i) a.swz[i] = 0;
    for (size_t i = 0; i < lhi.size(); ++i) a.swz[i] = (uint32_t)lhi[i] | ((uint32_t)mpos[i] << 8);
    for (int m = 0; m < (1 << K); ++m) {
        const uint32_t j = 64u * (uint32_t)m;
        a.off[m] = to_global(j);
        a.pm[m] = j ^ swizzle_bits(j, a.swz, a.nswz);
    }
    for (int s = 0; s < (1 << K); ++s) {
        uint32_t t = 0;
        for (int m = 0; m < K; ++m)
            if ((s >> (K - 1 - m)) & 1) t |= 1u << tpos[q[m]];
        a.ts[s] = t;
    }
    const uint64_t full = (n >= 64) ? ~0ULL : ((1ULL << n) - 1);
    a.rest = full & ~bmask;
    a.ntiles = 1ULL << (n - nb);
}

}  // namespace iqg
