// Split form of a tiled pass's one-qubit runs (host code, FP64): each run's
// 2x2 unitary as g diag(1, p) R diag(1, q) with R a real rotation, and the
// per-pass choice of which register group applies which diagonals.  Shared by
// the adjoint engine (adjoint.hip) and its CPU check (tests/split_check.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

namespace iqg {

// Entries of one phase table (tile bits 0-6, then 7-12).
constexpr int kPhaseTab = 192;

using split_cplx = std::complex<double>;

// u = g diag(1, p) R diag(1, q) with R = [[c, -s], [s, c]] real and |g| = |p| =
// |q| = 1 (Z-Y-Z angles, up to the global phase g), in FP64.  keep_q: a -1
// factor goes to p (diag(1, p) R diag(1, -1) = diag(1, -p) R(c, -s)), else to
// q, so that the side a pass leaves out of its tables is exactly 1 for
// rotations that have no diagonal there.  ok = false when the factors do not
// rebuild u to 1e-13.
struct SplitRun {
    split_cplx g, p, q;
    double c = 1.0, s = 0.0;
    bool ok = false;
};

inline SplitRun split_run(const split_cplx* um, bool keep_q) {
    SplitRun r;
    const split_cplx m00 = um[0], m01 = um[1], m10 = um[2], m11 = um[3];
    const double a = std::abs(m00), b = std::abs(m10), h = std::hypot(a, b);
    if (!(h > 0.5)) return r;
    r.c = a / h;
    r.s = b / h;
    r.g = a > 1e-300 ? m00 / a : split_cplx(1.0, 0.0);
    if (r.s >= r.c) {
        r.p = m10 / (r.g * r.s);
        r.q = -m01 / (r.g * r.s);
    } else if (r.s > 0.0) {
        r.p = m10 / (r.g * r.s);
        r.q = m11 / (r.g * r.p * r.c);
    } else {
        r.p = m11 / (r.g * r.c);
        r.q = 1.0;
    }
    r.p /= std::abs(r.p);
    r.q /= std::abs(r.q);
    const double snap = 1e-14;
    if (keep_q && std::abs(r.q + 1.0) < snap) {
        r.q = 1.0;
        r.p = -r.p;
        r.s = -r.s;
    }
    if (!keep_q && std::abs(r.p + 1.0) < snap) {  // diag(1, -1) R diag(1, q) = R(c, -s) diag(1, -q)
        r.p = 1.0;
        r.q = -r.q;
        r.s = -r.s;
    }
    if (std::abs(r.q - 1.0) < snap) r.q = 1.0;
    if (std::abs(r.p - 1.0) < snap) r.p = 1.0;
    const split_cplx w[4] = {r.g * r.c, -r.g * r.s * r.q, r.g * r.p * r.s, r.g * r.p * r.c * r.q};
    double err = 0.0;
    for (int i = 0; i < 4; ++i) err = std::max(err, std::abs(w[i] - um[i]));
    r.ok = err < 1e-13;
    return r;
}

// Split form of one tiled pass (runs ps[0, d.nseg) in group order; groups
// applied first to last: 0..ng-1 forward, ng-1..0 reverse).  Every run
// becomes g diag(1, p) R diag(1, q).  The first group applies R diag(1, q) of
// its runs and the q of all other runs as one phase per sub-cube; the last
// group diag(1, p) R and the other runs' p times the product of all g; the
// groups between, R alone (real: half the arithmetic of a complex 2x2).  Used
// when that costs fewer operations per element than a complex matrix per run
// (8 against 4 per real run, 4 per phase); the tables go to `tabs` (pairs).
template <typename Desc, typename Seg>
void split_pass(Desc& d, Seg* ps, bool reverse, std::vector<double>& tabs) {
    d.greal = 0;
    d.tpre = d.tpost = -1;
    const int ng = d.ngroups, nr = d.nseg;
    if (ng < 2 || d.gundo != 0 || d.ts_bits > 13) return;
    // groups in the order the sweep applies them: forward 0 .. ng-1; reverse
    // ng-1 .. 0, or with a register group (the last, d.greg) ng-2 .. 0 then it
    const int first_g = !reverse ? 0 : d.greg >= 0 ? ng - 2 : ng - 1;
    const int last_g = !reverse ? ng - 1 : d.greg >= 0 ? d.greg : 0;
    std::vector<SplitRun> sp(nr);
    std::vector<int> grp(nr);
    for (int gi = 0; gi < ng; ++gi)
        for (uint32_t k = 0; k < d.gsize[gi]; ++k) grp[d.gfirst[gi] + k] = gi;
    // forward: the ansatz's RZ RY has q = 1, so the pre table is not needed;
    // reverse: (RZ RY)^dagger has p = 1
    for (int r = 0; r < nr; ++r) {
        split_cplx u[4];
        for (int i = 0; i < 4; ++i) u[i] = split_cplx(ps[r].u[2 * i], ps[r].u[2 * i + 1]);
        sp[r] = split_run(u, !reverse);
        if (!sp[r].ok) return;
    }
    bool need_pre = false, need_post = false, first_real = true, last_real = true;
    split_cplx gall = 1.0;
    for (int r = 0; r < nr; ++r) {
        gall *= sp[r].g;
        if (grp[r] != first_g && sp[r].q != 1.0) need_pre = true;
        if (grp[r] != last_g && sp[r].p != 1.0) need_post = true;
        if (grp[r] == first_g && sp[r].q != 1.0) first_real = false;
        if (grp[r] == last_g && sp[r].p != 1.0) last_real = false;
    }
    // the product of the g: in the post table, else the pre table, else the
    // last group's first run
    const bool g_one = std::abs(gall - 1.0) < 1e-15;
    const bool g_in_run = !g_one && !need_post && !need_pre;
    const bool g_in_post = !g_one && need_post, g_in_pre = !g_one && !need_post && need_pre;
    if (g_in_run) last_real = false;
    int cost = 4 * (need_pre + need_post), full = 0;
    for (int r = 0; r < nr; ++r) {
        full += 8;
        const bool real = grp[r] == first_g ? first_real : grp[r] == last_g ? last_real : true;
        cost += real ? 4 : 8;
    }
    if (cost >= full) return;
    bool g_done = false;
    for (int r = 0; r < nr; ++r) {
        const SplitRun& x = sp[r];
        split_cplx m[4] = {x.c, -x.s, x.s, x.c};  // R
        if (grp[r] == first_g) {  // R diag(1, q)
            m[1] *= x.q;
            m[3] *= x.q;
        } else if (grp[r] == last_g) {  // diag(1, p) R
            m[2] *= x.p;
            m[3] *= x.p;
            if (g_in_run && !g_done) {
                for (auto& e : m) e *= gall;
                g_done = true;
            }
        }
        for (int i = 0; i < 4; ++i) {
            ps[r].v[2 * i] = m[i].real();
            ps[r].v[2 * i + 1] = m[i].imag();
        }
    }
    for (int gi = 0; gi < ng; ++gi)
        if (gi == first_g ? first_real : gi == last_g ? last_real : true) d.greal |= 1u << gi;
    // tables: entry i of the low part is the product over runs (outside the
    // group that applies it) on tile bits 0-6 set in i; the high part, bits
    // 7-12.  With fb[b] the product of those runs' factors on tile bit b, an
    // entry is the entry without its lowest set bit times that bit's fb: one
    // complex product per entry.
    auto table = [&](bool post) {
        const size_t at = tabs.size() / 2;
        tabs.resize(tabs.size() + 2 * kPhaseTab, 0.0);
        double* t = tabs.data() + 2 * at;
        split_cplx fb[13];
        for (auto& f : fb) f = 1.0;
        for (int r = 0; r < nr; ++r) {
            if (grp[r] == (post ? last_g : first_g)) continue;
            const int tb = ps[r].tbit;
            if (tb >= 0 && tb < 13) fb[tb] *= post ? sp[r].p : sp[r].q;
        }
        split_cplx e[kPhaseTab];
        e[0] = (post ? g_in_post : g_in_pre) ? gall : split_cplx(1.0);
        for (int i = 1; i < 128; ++i) e[i] = e[i & (i - 1)] * fb[__builtin_ctz((unsigned)i)];
        e[128] = 1.0;
        for (int i = 1; i < 64; ++i) e[128 + i] = e[128 + (i & (i - 1))] * fb[7 + __builtin_ctz((unsigned)i)];
        for (int i = 0; i < kPhaseTab; ++i) {
            t[2 * i] = e[i].real();
            t[2 * i + 1] = e[i].imag();
        }
        return (int)at;
    };
    if (need_pre) d.tpre = table(false);
    if (need_post) d.tpost = table(true);
}

}  // namespace iqg
