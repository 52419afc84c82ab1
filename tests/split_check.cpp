// CPU check of invariantq_amd/csrc/split.hpp.  A tiled pass's runs (one-qubit
// unitaries on distinct tile bits, in register groups of consecutive bits)
// are put in split form and applied the way the tile kernels apply them:
// groups in the sweep's order (forward 0..ng-1, reverse ng-1..0, or with a
// register group, the last, ng-2..0 and then it), the pre
// table's phase on each sub-cube before the first group's runs, each run's v
// (a real rotation where greal says so), the post table's phase after the
// last group's runs.  The result must equal every run's full matrix applied
// to the same random tile, to 1e-12, for Haar-random runs, the ansatz's RZ RY
// and its inverse, Pauli, Hadamard and phase gates, and products of these.
// Also checks that ansatz passes of three or more groups are split (the form
// saves work there).  Prints "ok <passes> <split>" or the first failure.
#include <algorithm>
#include <cstdio>
#include <random>

#include "split.hpp"

using iqg::split_cplx;

struct Seg {
    double u[8], v[8];
    int tbit, slot;
};
struct Desc {
    int ngroups = 0, nseg = 0, ts_bits = 0;
    uint32_t gundo = 0, gt0[16] = {}, gsize[16] = {}, gfirst[16] = {}, greal = 0;
    int tpre = -1, tpost = -1, greg = -1;
};

using M = split_cplx[4];

static void mul(const split_cplx* a, const split_cplx* b, split_cplx* o) {
    const split_cplx r[4] = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2],
                             a[2] * b[1] + a[3] * b[3]};
    for (int i = 0; i < 4; ++i) o[i] = r[i];
}

static void apply(std::vector<split_cplx>& x, int bit, const split_cplx* m) {
    for (size_t j = 0; j < x.size(); ++j)
        if (!((j >> bit) & 1u)) {
            const size_t k = j | (size_t(1) << bit);
            const split_cplx a = x[j], b = x[k];
            x[j] = m[0] * a + m[1] * b;
            x[k] = m[2] * a + m[3] * b;
        }
}

int main() {
    std::mt19937_64 rng(2026);
    std::uniform_real_distribution<double> ang(0.0, 2 * M_PI);
    std::normal_distribution<double> nd;
    const split_cplx I(0, 1);
    auto rz = [&](double t, split_cplx* m) { m[0] = std::exp(-I * (t / 2)); m[1] = m[2] = 0.0; m[3] = std::exp(I * (t / 2)); };
    auto ry = [&](double t, split_cplx* m) { m[0] = m[3] = std::cos(t / 2); m[1] = -std::sin(t / 2); m[2] = std::sin(t / 2); };
    auto haar = [&](split_cplx* m) {  // Gram-Schmidt of a Gaussian 2x2
        split_cplx a = {nd(rng), nd(rng)}, b = {nd(rng), nd(rng)}, c = {nd(rng), nd(rng)}, d = {nd(rng), nd(rng)};
        const double na = std::sqrt(std::norm(a) + std::norm(c));
        a /= na; c /= na;
        const split_cplx pr = std::conj(a) * b + std::conj(c) * d;
        b -= pr * a; d -= pr * c;
        const double nb = std::sqrt(std::norm(b) + std::norm(d));
        m[0] = a; m[1] = b / nb; m[2] = c; m[3] = d / nb;
    };
    auto fixed = [&](int k, split_cplx* m) {
        const double h = 1 / std::sqrt(2.0);
        const split_cplx tab[6][4] = {{0, 1, 1, 0}, {0, -I, I, 0}, {1, 0, 0, -1}, {h, h, h, -h},
                                      {1, 0, 0, I}, {1, 0, 0, std::exp(I * (M_PI / 4))}};
        for (int i = 0; i < 4; ++i) m[i] = tab[k][i];
    };
    int passes = 0, split = 0;
    for (int trial = 0; trial < 4000; ++trial) {
        const bool reverse = trial & 1;
        const int kind = (trial >> 1) % 4;  // 0 Haar, 1 ansatz, 2 fixed gates, 3 products
        const int ts = 6 + (int)(rng() % 8);  // 6..13 tile bits
        // runs on a random set of tile bits, cut into groups of <= 4 consecutive bits
        std::vector<int> bits;
        for (int b = 0; b < ts; ++b)
            if (rng() % 3) bits.push_back(b);
        if (bits.empty()) bits.push_back(0);
        Desc d;
        d.ts_bits = ts;
        std::vector<Seg> segs;
        for (size_t i = 0; i < bits.size();) {
            size_t j = i + 1;
            while (j < bits.size() && bits[j] == bits[j - 1] + 1 && j - i < 4) ++j;
            d.gt0[d.ngroups] = (uint32_t)bits[i];
            d.gsize[d.ngroups] = (uint32_t)(j - i);
            d.gfirst[d.ngroups] = (uint32_t)segs.size();
            ++d.ngroups;
            for (size_t k = i; k < j; ++k) {
                Seg s{};
                s.tbit = bits[k];
                M m;
                if (kind == 0) {
                    haar(m);
                } else if (kind == 1) {  // RZ(phi) RY(theta), or its inverse on the reverse sweep
                    M a, b;
                    rz(ang(rng), a);
                    ry(ang(rng), b);
                    if (!reverse) {
                        mul(a, b, m);
                    } else {
                        M t;
                        mul(a, b, t);
                        m[0] = std::conj(t[0]); m[1] = std::conj(t[2]); m[2] = std::conj(t[1]); m[3] = std::conj(t[3]);
                    }
                } else if (kind == 2) {
                    fixed((int)(rng() % 6), m);
                } else {
                    M a, b;
                    fixed((int)(rng() % 6), a);
                    haar(b);
                    mul(a, b, m);
                    if (rng() % 2) { rz(ang(rng), a); mul(a, m, m); }
                }
                for (int e = 0; e < 4; ++e) {
                    s.u[2 * e] = s.v[2 * e] = m[e].real();
                    s.u[2 * e + 1] = s.v[2 * e + 1] = m[e].imag();
                }
                segs.push_back(s);
            }
            i = j;
        }
        d.nseg = (int)segs.size();
        if (rng() % 2) d.greg = d.ngroups - 1;  // the last group applied in registers, last in both sweeps
        std::vector<double> tabs;
        iqg::split_pass(d, segs.data(), reverse, tabs);
        ++passes;
        if (d.tpre >= 0 || d.tpost >= 0 || d.greal) ++split;
        if (kind == 1 && d.ngroups >= 3 && !d.greal) {
            std::printf("ansatz pass of %d groups not split (trial %d)\n", d.ngroups, trial);
            return 1;
        }
        const size_t TS = size_t(1) << ts;
        std::vector<split_cplx> x(TS), y;
        for (auto& e : x) e = {nd(rng), nd(rng)};
        y = x;
        for (const Seg& s : segs) {
            M m;
            for (int e = 0; e < 4; ++e) m[e] = {s.u[2 * e], s.u[2 * e + 1]};
            apply(y, s.tbit, m);
        }
        auto phase = [&](int off, uint32_t base) {
            const double* t = tabs.data() + 2 * off;
            const split_cplx a(t[2 * (base & 127u)], t[2 * (base & 127u) + 1]);
            const uint32_t h = 128u + (base >> 7);
            return a * split_cplx(t[2 * h], t[2 * h + 1]);
        };
        std::vector<int> order;
        for (int k = 0; k < d.ngroups; ++k) order.push_back(k);
        if (reverse) {
            std::reverse(order.begin(), order.end());
            if (d.greg >= 0) {  // ng-2 .. 0, then the register group
                order.erase(order.begin());
                order.push_back(d.greg);
            }
        }
        const int first_g = order.front(), last_g = order.back();
        for (int g : order) {
            const uint32_t gmask = ((1u << d.gsize[g]) - 1) << d.gt0[g];
            for (uint32_t base = 0; base < TS; ++base) {
                if (base & gmask) continue;
                if (g == first_g && d.tpre >= 0) {
                    const split_cplx ph = phase(d.tpre, base);
                    for (uint32_t m = 0; m < (1u << d.gsize[g]); ++m) x[base | (m << d.gt0[g])] *= ph;
                }
            }
            for (uint32_t i = 0; i < d.gsize[g]; ++i) {
                const Seg& s = segs[d.gfirst[g] + i];
                M m;
                if ((d.greal >> g) & 1u) {
                    m[0] = m[3] = s.v[0];
                    m[1] = -s.v[4];
                    m[2] = s.v[4];
                    if (s.v[1] != 0 || s.v[3] != 0 || s.v[5] != 0 || s.v[7] != 0 || s.v[2] != -s.v[4] || s.v[6] != s.v[0]) {
                        std::printf("group %d marked real but run %u has v = (%g,%g,%g,%g,%g,%g,%g,%g)\n", g, i, s.v[0],
                                    s.v[1], s.v[2], s.v[3], s.v[4], s.v[5], s.v[6], s.v[7]);
                        return 1;
                    }
                } else {
                    for (int e = 0; e < 4; ++e) m[e] = {s.v[2 * e], s.v[2 * e + 1]};
                }
                apply(x, s.tbit, m);
            }
            for (uint32_t base = 0; base < TS; ++base) {
                if (base & gmask) continue;
                if (g == last_g && d.tpost >= 0) {
                    const split_cplx ph = phase(d.tpost, base);
                    for (uint32_t m = 0; m < (1u << d.gsize[g]); ++m) x[base | (m << d.gt0[g])] *= ph;
                }
            }
        }
        double err = 0.0;
        for (size_t j = 0; j < TS; ++j) err = std::max(err, std::abs(x[j] - y[j]));
        if (!(err < 1e-12)) {
            std::printf("trial %d (kind %d, reverse %d, ts %d, groups %d, greal %x, pre %d, post %d): error %.3e\n",
                        trial, kind, (int)reverse, ts, d.ngroups, d.greal, d.tpre, d.tpost, err);
            return 1;
        }
    }
    std::printf("ok %d %d\n", passes, split);
    return 0;
}
