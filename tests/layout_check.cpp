// Host-only check of the gate kernels' index layout (invariantq_amd/csrc/layout.hpp),
// built with g++ by tests/test_layout.py.  For every target list given on stdin
// ("n K q0 .. qK-1" per line) it replays the kernels' address arithmetic and
// verifies, without touching a GPU:
//   direct: every group member index is < 2^n, each amplitude belongs to exactly
//           one (group, member) pair, and member s has the MSB-first target bits of s;
//   tiled:  every global address is < 2^n and each amplitude is loaded exactly once;
//           the LDS positions of a tile are a permutation of [0, 64*2^K);
//           the element a lane gathers for local index s is the amplitude whose
//           target bits spell s and whose other bits are that lane's group;
//           16 consecutive lanes hit 16 distinct 16-element slots in both passes;
//   mfma:   each lane's loads and stores stay inside its group, read/write every
//           amplitude exactly once, and land on the local indices the matrix
//           fragments were permuted to.
// Prints "ok" or the first violation.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "layout.hpp"

using namespace iqg;

static int fail(const char* what, int n, int K) {
    std::printf("FAIL %s (n=%d K=%d)\n", what, n, K);
    return 1;
}

template <int K>
static int check(int n, const uint32_t* q) {
    const uint64_t N = 1ULL << n;
    const int D = 1 << K;
    auto local_of = [&](uint64_t idx) {
        uint32_t l = 0;
        for (int m = 0; m < K; ++m) l |= (uint32_t)((idx >> q[m]) & 1ULL) << (K - 1 - m);
        return l;
    };
    uint64_t tmask = 0;
    for (int m = 0; m < K; ++m) tmask |= 1ULL << q[m];
    {  // direct
        DirectArgs<K> a;
        fill_direct_args<K>(a, q);
        std::vector<unsigned char> seen(N, 0);
        for (uint64_t g = 0; g < (N >> K); ++g) {
            const uint64_t base = insert_zero_bits<K>(g, a.pos);
            if (base & tmask) return fail("direct base has target bits", n, K);
            for (int s = 0; s < D; ++s) {
                const uint64_t idx = base + a.off[s];
                if (idx >= N) return fail("direct index out of range", n, K);
                if (seen[idx]++) return fail("direct index repeated", n, K);
                if (local_of(idx) != (uint32_t)s) return fail("direct local index", n, K);
            }
        }
    }
    for (int f64 = 0; f64 < 2 && K == 5 && n >= 9; ++f64) {  // matrix-core mapping, both row layouts
        MfmaArgs ma;
        uint32_t s_in[32], s_out[32];
        fill_mfma_args(ma, n, q, f64 == 1, s_in, s_out);
        std::set<uint32_t> pin(s_in, s_in + 32), pout(s_out, s_out + 32);
        if (pin.size() != 32 || pout.size() != 32) return fail("mfma local maps not permutations", n, K);
        // the kernel's traversal: super-batch sb0 -> rotate_index -> UB batches;
        // every rotation must visit each batch exactly once
        std::vector<uint64_t> order;
        for (uint32_t ub : {1u, 2u, 4u})
            for (uint32_t rot : {0u, 1u, 2u, 5u}) {
                if (ma.nbatches % ub) continue;  // the launcher lowers UB until it divides
                const uint64_t nsuper = ma.nbatches / ub;
                uint32_t bits = 0;
                while ((1ULL << bits) < nsuper) ++bits;
                if ((1ULL << bits) != nsuper) return fail("mfma super-batch count not a power of two", n, K);
                std::vector<unsigned char> hit(ma.nbatches, 0);
                std::vector<uint64_t> ord;
                for (uint64_t sb0 = 0; sb0 < nsuper; ++sb0) {
                    const uint64_t sb = rotate_index(sb0, rot, bits);
                    if (sb >= nsuper) return fail("mfma rotated super-batch out of range", n, K);
                    for (uint32_t u = 0; u < ub; ++u) {
                        const uint64_t b = sb * ub + u;
                        if (hit[b]++) return fail("mfma batch visited twice", n, K);
                        ord.push_back(b);
                    }
                }
                if (ord.size() != ma.nbatches) return fail("mfma batch missed", n, K);
                if (rot == 2 || order.empty()) order = ord;
            }
        std::vector<unsigned char> rd(N, 0), wr(N, 0);
        for (uint64_t b : order)
            for (uint32_t lane = 0; lane < 64; ++lane) {
                const uint32_t qq = lane >> 4, nn = lane & 15;
                uint32_t pos5[5];
                for (int m = 0; m < 5; ++m) pos5[m] = ma.pos[m];
                const uint64_t base = insert_zero_bits<5>(b * 16 + nn, pos5);
                for (int kb = 0; kb < 8; ++kb) {
                    const uint64_t idx = base + ma.xq[qq] + ma.xk[kb];
                    if (idx >= N) return fail("mfma load out of range", n, K);
                    if (rd[idx]++) return fail("mfma load repeated", n, K);
                    if (local_of(idx) != s_in[4 * kb + qq]) return fail("mfma load local index", n, K);
                    if ((idx & ~tmask) != base) return fail("mfma load leaves its group", n, K);
                }
                for (int mt = 0; mt < 2; ++mt)
                    for (int jj = 0; jj < 4; ++jj) {
                        const uint64_t idx = base + ma.yq[qq] + ma.yk[4 * mt + jj];
                        if (idx >= N) return fail("mfma store out of range", n, K);
                        if (wr[idx]++) return fail("mfma store repeated", n, K);
                        if (local_of(idx) != s_out[mfma_row(qq, jj, mt, f64 == 1)]) return fail("mfma store local index", n, K);
                        if ((idx & ~tmask) != base) return fail("mfma store leaves its group", n, K);
                    }
            }
        for (uint64_t i = 0; i < N; ++i)
            if (rd[i] != 1 || wr[i] != 1) return fail("mfma pass missed an amplitude", n, K);
    }
    if (n < K + 6) return 0;
    TiledArgs<K> a;
    fill_tiled_args<K>(a, n, q);
    const int T = 64 * D;
    std::vector<unsigned char> seen(N, 0);
    for (uint64_t t = 0; t < a.ntiles; ++t) {
        std::vector<uint64_t> lds(T, ~0ULL);  // LDS position -> global index
        for (uint32_t lane = 0; lane < 64; ++lane) {
            uint32_t pl, pg;
            tile_lane_positions<K>(lane, a, pl, pg);
            const uint64_t base = deposit_bits(t, a.rest) + lane;
            for (int m = 0; m < D; ++m) {
                const uint64_t gidx = base + a.off[m];
                const uint32_t pos = pl ^ a.pm[m];
                if (gidx >= N) return fail("tiled global index out of range", n, K);
                if (pos >= (uint32_t)T) return fail("tiled LDS position out of range", n, K);
                if (seen[gidx]++) return fail("tiled global index repeated", n, K);
                if (lds[pos] != ~0ULL) return fail("tiled LDS position repeated", n, K);
                lds[pos] = gidx;
            }
        }
        std::set<uint64_t> groups;
        for (uint32_t lane = 0; lane < 64; ++lane) {
            uint32_t pl, pg;
            tile_lane_positions<K>(lane, a, pl, pg);
            uint64_t rest_bits = ~0ULL;
            for (int s = 0; s < D; ++s) {
                const uint32_t pos = pg ^ a.ts[s];
                if (pos >= (uint32_t)T) return fail("gather position out of range", n, K);
                const uint64_t gidx = lds[pos];
                if (local_of(gidx) != (uint32_t)s) return fail("gather local index", n, K);
                const uint64_t r = gidx & ~tmask;
                if (s == 0) rest_bits = r;
                else if (r != rest_bits) return fail("gather mixes groups", n, K);
            }
            if (!groups.insert(rest_bits).second) return fail("two lanes share a group", n, K);
        }
        // bank slots (16 elements of 16 B = 256 B) for 16-lane phases
        for (uint32_t l0 = 0; l0 < 64; l0 += 16) {
            for (int m = 0; m < D; ++m) {
                std::set<uint32_t> slots;
                for (uint32_t lane = l0; lane < l0 + 16; ++lane) {
                    uint32_t pl, pg;
                    tile_lane_positions<K>(lane, a, pl, pg);
                    slots.insert((pl ^ a.pm[m]) & 15u);
                }
                if (slots.size() != 16) return fail("coalesced pass bank conflict", n, K);
            }
            for (int s = 0; s < D; ++s) {
                std::set<uint32_t> slots;
                for (uint32_t lane = l0; lane < l0 + 16; ++lane) {
                    uint32_t pl, pg;
                    tile_lane_positions<K>(lane, a, pl, pg);
                    slots.insert((pg ^ a.ts[s]) & 15u);
                }
                if (slots.size() != 16) return fail("gather pass bank conflict", n, K);
            }
        }
    }
    for (uint64_t i = 0; i < N; ++i)
        if (seen[i] != 1) return fail("tiled pass missed an amplitude", n, K);
    return 0;
}

int main() {
    int n, K, cases = 0;
    while (std::scanf("%d %d", &n, &K) == 2) {
        uint32_t q[8];
        for (int m = 0; m < K; ++m)
            if (std::scanf("%u", &q[m]) != 1) return 1;
        int r = 0;
        switch (K) {
            case 1: r = check<1>(n, q); break;
            case 2: r = check<2>(n, q); break;
            case 3: r = check<3>(n, q); break;
            case 4: r = check<4>(n, q); break;
            case 5: r = check<5>(n, q); break;
            default: std::printf("FAIL bad K\n"); return 1;
        }
        if (r) {
            std::printf("case n=%d K=%d q=", n, K);
            for (int m = 0; m < K; ++m) std::printf("%u ", q[m]);
            std::printf("\n");
            return 1;
        }
        ++cases;
    }
    std::printf("ok %d\n", cases);
    return 0;
}
