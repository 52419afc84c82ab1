// TEST INFRASTRUCTURE ONLY — the oracle.
//
// A CPU (C++17/OpenMP) restatement of the reference InvariantQ native core's
// statevector and density-matrix algorithms.  It is the parity checker for the
// HIP engine and the timed CPU baseline.  It is never linked into, loaded by or
// called from the product library (invariantq_amd/core/libinvariantq.so); only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// Every routine cites the reference file:line whose algorithm it restates.
// Loop structure and `schedule(static)` follow the reference so that the
// timing is a fair baseline; two documented deviations:
//   * apply_fused uses the documented MSB-first local index for every k
//     (reference core/src/statevector.cpp:133-139 is LSB-first for k=3,4; the
//     two agree after reversing the qubit list), and the k<=4 guard
//     (statevector.cpp:99-102) is lifted to k<=5 for 32x32 blocks;
//   * errors are returned as codes instead of C++ exceptions crossing extern "C".
//
// Pinning: tests/test_oracle_*.py check every function against closed forms,
// brute-force Kronecker products, invariants, and the reference core compiled
// from source (oracle/_ref) through the committed fixtures in tests/golden/.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include <omp.h>

namespace iqo {

using cplx = std::complex<double>;

static thread_local int g_last_error = 0;  // 0 = ok, 1 = invalid argument, 2 = zero-probability collapse

struct SV {
    size_t n;
    uint64_t dim;
    std::vector<cplx> amp;
    explicit SV(size_t nq) : n(nq), dim(1ULL << nq), amp(dim, cplx(0.0, 0.0)) { amp[0] = 1.0; }
};

struct DM {
    size_t n;
    uint64_t dim;
    std::vector<cplx> rho;  // row-major rho[r*dim + c] (reference density_matrix.cpp:9)
    explicit DM(size_t nq) : n(nq), dim(1ULL << nq), rho(dim * dim, cplx(0.0, 0.0)) { rho[0] = 1.0; }
};

// ---------------------------------------------------------------- statevector

// reference core/src/statevector.cpp:26-47 — pair (i0, i0|2^t) from a half-index.
static void sv_apply_1q(SV& s, size_t t, const cplx* U) {
    const uint64_t bit = 1ULL << t;
    const uint64_t lo = bit - 1;
    const uint64_t npairs = s.dim >> 1;
    cplx* A = s.amp.data();
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < npairs; ++p) {
        const uint64_t i0 = ((p & ~lo) << 1) | (p & lo);
        const uint64_t i1 = i0 | bit;
        const cplx x0 = A[i0];
        const cplx x1 = A[i1];
        A[i0] = U[0] * x0 + U[1] * x1;
        A[i1] = U[2] * x0 + U[3] * x1;
    }
}

// reference core/src/statevector.cpp:49-87 — quad of indices; local row index 2*b(q0)+b(q1).
static void sv_apply_2q(SV& s, size_t q0, size_t q1, const cplx* U) {
    const size_t ql = std::min(q0, q1), qh = std::max(q0, q1);
    const uint64_t nquads = s.dim >> 2;
    const uint64_t b0 = 1ULL << q0, b1 = 1ULL << q1;
    cplx* A = s.amp.data();
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < nquads; ++p) {
        const uint64_t low = p & ((1ULL << ql) - 1);
        const uint64_t mid = (p >> ql) & ((1ULL << (qh - ql - 1)) - 1);
        const uint64_t high = p >> (qh - 1);
        const uint64_t base = (high << (qh + 1)) | (mid << (ql + 1)) | low;
        const uint64_t idx[4] = {base, base | b1, base | b0, base | b0 | b1};
        cplx v[4];
        for (int k = 0; k < 4; ++k) v[k] = A[idx[k]];
        for (int r = 0; r < 4; ++r)
            A[idx[r]] = U[4 * r + 0] * v[0] + U[4 * r + 1] * v[1] + U[4 * r + 2] * v[2] + U[4 * r + 3] * v[3];
    }
}

// reference core/src/statevector.cpp:89-152 — gather 2^k amplitudes, dense matvec,
// scatter.  The outer index is spread over the non-target qubits in ascending
// order (statevector.cpp:114-128).  Local index is MSB-first: qubits[0] is the
// most significant local bit.  Per-iteration std::vector gathers are kept
// (statevector.cpp:130-131) so the baseline has the reference's cost model.
static void sv_apply_fused(SV& s, const std::vector<size_t>& q, const cplx* U) {
    const size_t k = q.size();
    if (k == 1) { sv_apply_1q(s, q[0], U); return; }
    if (k == 2) { sv_apply_2q(s, q[0], q[1], U); return; }
    const uint64_t kd = 1ULL << k;
    const uint64_t outer = s.dim >> k;
    std::vector<uint64_t> lmask(k);
    for (size_t j = 0; j < k; ++j) lmask[j] = 1ULL << q[k - 1 - j];  // local bit j <-> qubits[k-1-j]
    std::vector<size_t> others;
    for (size_t b = 0; b < s.n; ++b)
        if (std::find(q.begin(), q.end(), b) == q.end()) others.push_back(b);
    cplx* A = s.amp.data();
#pragma omp parallel for schedule(static)
    for (uint64_t o = 0; o < outer; ++o) {
        uint64_t base = 0;
        for (size_t b = 0; b < others.size(); ++b)
            if ((o >> b) & 1ULL) base |= 1ULL << others[b];
        std::vector<uint64_t> idx(kd);
        std::vector<cplx> v(kd);
        for (uint64_t sidx = 0; sidx < kd; ++sidx) {
            uint64_t g = base;
            for (size_t j = 0; j < k; ++j)
                if ((sidx >> j) & 1ULL) g |= lmask[j];
            idx[sidx] = g;
            v[sidx] = A[g];
        }
        for (uint64_t r = 0; r < kd; ++r) {
            cplx acc(0.0, 0.0);
            for (uint64_t c = 0; c < kd; ++c) acc += U[r * kd + c] * v[c];
            A[idx[r]] = acc;
        }
    }
}

// reference core/src/statevector.cpp:154-169
static double sv_measure_prob(const SV& s, size_t t, int outcome) {
    const uint64_t bit = 1ULL << t;
    double p = 0.0;
    const cplx* A = s.amp.data();
#pragma omp parallel for reduction(+ : p) schedule(static)
    for (uint64_t i = 0; i < s.dim; ++i) {
        const int b = (i & bit) ? 1 : 0;
        if (b == outcome) p += std::norm(A[i]);
    }
    return p;
}

// reference core/src/statevector.cpp:171-188; returns false for p < 1e-15.
static bool sv_collapse(SV& s, size_t t, int outcome) {
    const double p = sv_measure_prob(s, t, outcome);
    if (p < 1e-15) return false;
    const double f = 1.0 / std::sqrt(p);
    const uint64_t bit = 1ULL << t;
    cplx* A = s.amp.data();
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < s.dim; ++i) {
        const int b = (i & bit) ? 1 : 0;
        if (b == outcome) A[i] *= f;
        else A[i] = cplx(0.0, 0.0);
    }
    return true;
}

// reference core/src/statevector.cpp:206-213
static double sv_norm2(const SV& s) {
    double acc = 0.0;
    const cplx* A = s.amp.data();
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (uint64_t i = 0; i < s.dim; ++i) acc += std::norm(A[i]);
    return acc;
}

// reference core/src/statevector.cpp:215-224: scale by 1/sqrt(norm^2) when it is positive
static void sv_normalize(SV& s) {
    const double n2 = sv_norm2(s);
    if (n2 > 0.0) {
        const double f = 1.0 / std::sqrt(n2);
        cplx* A = s.amp.data();
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < s.dim; ++i) A[i] *= f;
    }
}

// reference core/src/statevector.cpp:226-233
static void sv_probs(const SV& s, double* out) {
    const cplx* A = s.amp.data();
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < s.dim; ++i) out[i] = std::norm(A[i]);
}

// reference core/src/statevector.cpp:235-256 — serial CDF, mt19937_64(seed),
// uniform_real_distribution<double>(0,1), lower_bound, clamp to dim-1.
static void sv_sample(const SV& s, size_t shots, uint64_t seed, uint64_t* out) {
    std::vector<double> p(s.dim);
    sv_probs(s, p.data());
    std::vector<double> cdf(s.dim);
    double run = 0.0;
    for (uint64_t i = 0; i < s.dim; ++i) { run += p[i]; cdf[i] = run; }
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    for (size_t k = 0; k < shots; ++k) {
        const double u = uni(gen);
        uint64_t j = (uint64_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (j >= s.dim) j = s.dim - 1;
        out[k] = j;
    }
}

// reference core/src/statevector.cpp:258-291 — serial; skips |a_i|^2 < 1e-18;
// characters other than X/Y/Z act as identity.
static cplx sv_expval_pauli(const SV& s, const char* ops, const size_t* q, size_t len) {
    cplx acc(0.0, 0.0);
    const cplx* A = s.amp.data();
    for (uint64_t i = 0; i < s.dim; ++i) {
        const cplx ai = A[i];
        if (std::norm(ai) < 1e-18) continue;
        uint64_t j = i;
        cplx ph(1.0, 0.0);
        for (size_t t = 0; t < len; ++t) {
            const uint64_t b = (i >> q[t]) & 1ULL;
            switch (ops[t]) {
                case 'X': j ^= 1ULL << q[t]; break;
                case 'Y': j ^= 1ULL << q[t]; ph *= (b == 0) ? cplx(0, 1) : cplx(0, -1); break;
                case 'Z': if (b) ph *= -1.0; break;
                default: break;
            }
        }
        acc += std::conj(A[j]) * ph * ai;
    }
    return acc;
}

// reference core/src/statevector.cpp:293-330 — reduced density matrix of the
// kept qubits (local bit b <-> keep[b], LSB-first, as the reference).
static std::vector<cplx> sv_partial_trace(const SV& s, const std::vector<size_t>& keep) {
    const size_t k = keep.size();
    const uint64_t kd = 1ULL << k;
    std::vector<size_t> tr;
    for (size_t b = 0; b < s.n; ++b)
        if (std::find(keep.begin(), keep.end(), b) == keep.end()) tr.push_back(b);
    const uint64_t td = 1ULL << tr.size();
    std::vector<cplx> red(kd * kd, cplx(0.0, 0.0));
    for (uint64_t r = 0; r < kd; ++r)
        for (uint64_t c = 0; c < kd; ++c) {
            uint64_t ir = 0, ic = 0;
            for (size_t b = 0; b < k; ++b) {
                if ((r >> b) & 1ULL) ir |= 1ULL << keep[b];
                if ((c >> b) & 1ULL) ic |= 1ULL << keep[b];
            }
            cplx sum(0.0, 0.0);
            for (uint64_t e = 0; e < td; ++e) {
                uint64_t m = 0;
                for (size_t b = 0; b < tr.size(); ++b)
                    if ((e >> b) & 1ULL) m |= 1ULL << tr[b];
                sum += s.amp[ir | m] * std::conj(s.amp[ic | m]);
            }
            red[r * kd + c] = sum;
        }
    return red;
}

// reference core/src/statevector.cpp:333-393 (and density_matrix.cpp:322-370):
// largest-pivot complex Jacobi with an iteration cap and an off-diagonal threshold.
static std::vector<double> jacobi_eigenvalues(const std::vector<cplx>& m, size_t d, int max_iter, double tol) {
    std::vector<cplx> A(m);
    auto at = [&](size_t i, size_t j) -> cplx& { return A[i * d + j]; };
    for (int it = 0; it < max_iter; ++it) {
        double best = 0.0;
        size_t p = 0, q = 1;
        for (size_t i = 0; i < d; ++i)
            for (size_t j = i + 1; j < d; ++j) {
                const double v = std::abs(at(i, j));
                if (v > best) { best = v; p = i; q = j; }
            }
        if (best < tol) break;
        const double app = at(p, p).real(), aqq = at(q, q).real();
        const cplx apq = at(p, q);
        const double th = 0.5 * std::atan2(2.0 * std::abs(apq), app - aqq);
        const cplx ph = apq / std::abs(apq);
        const double c = std::cos(th);
        const cplx sn = std::sin(th) * ph;
        for (size_t i = 0; i < d; ++i) {
            if (i == p || i == q) continue;
            const cplx aip = at(i, p), aiq = at(i, q);
            at(i, p) = c * aip + std::conj(sn) * aiq;
            at(p, i) = std::conj(at(i, p));
            at(i, q) = -sn * aip + c * aiq;
            at(q, i) = std::conj(at(i, q));
        }
        const double cross = 2.0 * (c * std::conj(sn) * apq).real();
        const double npp = c * c * app + std::norm(sn) * aqq + cross;
        const double nqq = std::norm(sn) * app + c * c * aqq - cross;
        at(p, p) = npp;
        at(q, q) = nqq;
        at(p, q) = 0.0;
        at(q, p) = 0.0;
    }
    std::vector<double> ev(d);
    for (size_t i = 0; i < d; ++i) ev[i] = at(i, i).real();
    return ev;
}

// reference core/src/statevector.cpp:395-410
static double sv_entropy(const SV& s, const std::vector<size_t>& sub) {
    const size_t k = sub.size();
    if (k == 0 || k == s.n) return 0.0;
    const auto red = sv_partial_trace(s, sub);
    const auto ev = jacobi_eigenvalues(red, 1ULL << k, 100, 1e-12);
    double h = 0.0;
    for (double l : ev)
        if (l > 1e-14) h -= l * std::log2(l);
    return std::max(0.0, h);
}

// --------------------------------------------------------------- density matrix

// reference core/src/density_matrix.cpp:18-67 — left multiply into a fresh zeroed
// buffer (column-strided walk), then right multiply by U^dagger back into rho.
static void dm_apply_1q(DM& d, size_t t, const cplx* U) {
    const uint64_t bit = 1ULL << t, lo = bit - 1, half = d.dim >> 1, D = d.dim;
    std::vector<cplx> nxt(D * D, cplx(0.0, 0.0));
    cplx* R = d.rho.data();
#pragma omp parallel for schedule(static)
    for (uint64_t c = 0; c < D; ++c)
        for (uint64_t p = 0; p < half; ++p) {
            const uint64_t r0 = ((p & ~lo) << 1) | (p & lo), r1 = r0 | bit;
            const cplx x0 = R[r0 * D + c], x1 = R[r1 * D + c];
            nxt[r0 * D + c] = U[0] * x0 + U[1] * x1;
            nxt[r1 * D + c] = U[2] * x0 + U[3] * x1;
        }
    const cplx Ud[4] = {std::conj(U[0]), std::conj(U[2]), std::conj(U[1]), std::conj(U[3])};
#pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < D; ++r)
        for (uint64_t p = 0; p < half; ++p) {
            const uint64_t c0 = ((p & ~lo) << 1) | (p & lo), c1 = c0 | bit;
            const cplx x0 = nxt[r * D + c0], x1 = nxt[r * D + c1];
            R[r * D + c0] = x0 * Ud[0] + x1 * Ud[2];
            R[r * D + c1] = x0 * Ud[1] + x1 * Ud[3];
        }
}

// reference core/src/density_matrix.cpp:69-145
static void dm_apply_2q(DM& d, size_t q0, size_t q1, const cplx* U) {
    const size_t ql = std::min(q0, q1), qh = std::max(q0, q1);
    const uint64_t D = d.dim, quarter = D >> 2, b0 = 1ULL << q0, b1 = 1ULL << q1;
    auto quad = [&](uint64_t p, uint64_t* idx) {
        const uint64_t low = p & ((1ULL << ql) - 1);
        const uint64_t mid = (p >> ql) & ((1ULL << (qh - ql - 1)) - 1);
        const uint64_t high = p >> (qh - 1);
        const uint64_t base = (high << (qh + 1)) | (mid << (ql + 1)) | low;
        idx[0] = base; idx[1] = base | b1; idx[2] = base | b0; idx[3] = base | b0 | b1;
    };
    std::vector<cplx> nxt(D * D, cplx(0.0, 0.0));
    cplx* R = d.rho.data();
#pragma omp parallel for schedule(static)
    for (uint64_t c = 0; c < D; ++c)
        for (uint64_t p = 0; p < quarter; ++p) {
            uint64_t idx[4];
            quad(p, idx);
            cplx v[4];
            for (int k = 0; k < 4; ++k) v[k] = R[idx[k] * D + c];
            for (int r = 0; r < 4; ++r)
                nxt[idx[r] * D + c] = U[4 * r] * v[0] + U[4 * r + 1] * v[1] + U[4 * r + 2] * v[2] + U[4 * r + 3] * v[3];
        }
    cplx Ud[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) Ud[4 * r + c] = std::conj(U[4 * c + r]);
#pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < D; ++r)
        for (uint64_t p = 0; p < quarter; ++p) {
            uint64_t idx[4];
            quad(p, idx);
            cplx v[4];
            for (int k = 0; k < 4; ++k) v[k] = nxt[r * D + idx[k]];
            for (int c = 0; c < 4; ++c)
                R[r * D + idx[c]] = v[0] * Ud[c] + v[1] * Ud[4 + c] + v[2] * Ud[8 + c] + v[3] * Ud[12 + c];
        }
}

// reference core/src/density_matrix.cpp:147-163 — one full copy per Kraus
// operator, unitary-style conjugation, serial accumulate.
static void dm_kraus_1q(DM& d, size_t t, const std::vector<cplx>& ks) {
    const size_t nk = ks.size() / 4;
    const uint64_t total = d.dim * d.dim;
    std::vector<cplx> acc(total, cplx(0.0, 0.0));
    const std::vector<cplx> start = d.rho;
    for (size_t k = 0; k < nk; ++k) {
        d.rho = start;
        dm_apply_1q(d, t, &ks[4 * k]);
        for (uint64_t i = 0; i < total; ++i) acc[i] += d.rho[i];
    }
    d.rho = std::move(acc);
}

// reference core/src/density_matrix.cpp:193-199
static double dm_trace(const DM& d) {
    cplx tr(0.0, 0.0);
    for (uint64_t i = 0; i < d.dim; ++i) tr += d.rho[i * d.dim + i];
    return tr.real();
}

// reference core/src/density_matrix.cpp:201-208
static double dm_purity(const DM& d) {
    double acc = 0.0;
    const uint64_t total = d.dim * d.dim;
    const cplx* R = d.rho.data();
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (uint64_t i = 0; i < total; ++i) acc += std::norm(R[i]);
    return acc;
}

// reference core/src/density_matrix.cpp:218-229
static double dm_measure_prob(const DM& d, size_t t, int outcome) {
    const uint64_t bit = 1ULL << t;
    double p = 0.0;
    for (uint64_t i = 0; i < d.dim; ++i) {
        const int b = (i & bit) ? 1 : 0;
        if (b == outcome) p += d.rho[i * d.dim + i].real();
    }
    return std::max(0.0, p);
}

// reference core/src/density_matrix.cpp:231-256
static bool dm_collapse(DM& d, size_t t, int outcome) {
    const double p = dm_measure_prob(d, t, outcome);
    if (p < 1e-15) return false;
    const double f = 1.0 / p;
    const uint64_t bit = 1ULL << t, D = d.dim;
    cplx* R = d.rho.data();
#pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < D; ++r) {
        const int br = (r & bit) ? 1 : 0;
        for (uint64_t c = 0; c < D; ++c) {
            const int bc = (c & bit) ? 1 : 0;
            if (br == outcome && bc == outcome) R[r * D + c] *= f;
            else R[r * D + c] = cplx(0.0, 0.0);
        }
    }
    return true;
}

// reference core/src/density_matrix.cpp:310-380 — diagonal Shannon entropy for
// dim > 64, capped Jacobi otherwise.
static double dm_entropy(const DM& d) {
    if (d.dim > 64) {
        double h = 0.0;
        for (uint64_t i = 0; i < d.dim; ++i) {
            const double p = std::max(0.0, d.rho[i * d.dim + i].real());
            if (p > 1e-14) h -= p * std::log2(p);
        }
        return h;
    }
    const auto ev = jacobi_eigenvalues(d.rho, d.dim, 120, 1e-13);
    double h = 0.0;
    for (double l : ev)
        if (l > 1e-14) h -= l * std::log2(l);
    return std::max(0.0, h);
}

static std::vector<cplx> pack(const double* re, const double* im, size_t count) {
    std::vector<cplx> v(count);
    for (size_t i = 0; i < count; ++i) v[i] = cplx(re[i], im[i]);
    return v;
}

}  // namespace iqo

using namespace iqo;

// The C names follow reference core/include/c_api.h:10-38 so that tests can
// drive the oracle and the HIP engine through one binding shape.
extern "C" {

int iqo_last_error(void) { return g_last_error; }

// The capped largest-pivot Jacobi above on a caller's Hermitian matrix
// (d x d row-major, split re/im): the reference's eigenvalue method, exposed
// so faster restatements can be checked against it bit for bit.
void iqo_hermitian_eigenvalues(const double* re, const double* im, size_t d, int max_iter, double tol,
                               double* out) {
    std::vector<cplx> m(d * d);
    for (size_t i = 0; i < d * d; ++i) m[i] = cplx(re[i], im[i]);
    const auto ev = jacobi_eigenvalues(m, d, max_iter, tol);
    for (size_t i = 0; i < d; ++i) out[i] = ev[i];
}

void* iq_statevector_create(size_t n) {
    if (n > 40) return nullptr;
    try { return new SV(n); } catch (...) { return nullptr; }
}
void iq_statevector_free(void* p) { delete static_cast<SV*>(p); }
void iq_statevector_reset(void* p) {
    if (!p) return;
    SV& s = *static_cast<SV*>(p);
    std::fill(s.amp.begin(), s.amp.end(), cplx(0.0, 0.0));
    s.amp[0] = 1.0;
}
void iq_statevector_apply_1q(void* p, size_t t, const double* re, const double* im) {
    if (!p || !re || !im) return;
    SV& s = *static_cast<SV*>(p);
    if (t >= s.n) { g_last_error = 1; return; }
    const auto U = pack(re, im, 4);
    sv_apply_1q(s, t, U.data());
}
void iq_statevector_apply_2q(void* p, size_t q0, size_t q1, const double* re, const double* im) {
    if (!p || !re || !im) return;
    SV& s = *static_cast<SV*>(p);
    if (q0 >= s.n || q1 >= s.n || q0 == q1) { g_last_error = 1; return; }
    const auto U = pack(re, im, 16);
    sv_apply_2q(s, q0, q1, U.data());
}
void iq_statevector_apply_fused(void* p, const size_t* qubits, size_t k, const double* re, const double* im) {
    if (!p || !qubits || !re || !im) return;
    SV& s = *static_cast<SV*>(p);
    if (k == 0 || k > 5 || k > s.n) { g_last_error = 1; return; }
    std::vector<size_t> q(qubits, qubits + k);
    for (size_t i = 0; i < k; ++i) {
        if (q[i] >= s.n) { g_last_error = 1; return; }
        for (size_t j = 0; j < i; ++j)
            if (q[i] == q[j]) { g_last_error = 1; return; }
    }
    const auto U = pack(re, im, (1ULL << k) * (1ULL << k));
    sv_apply_fused(s, q, U.data());
}
double iq_statevector_measure_prob(void* p, size_t t, int outcome) {
    if (!p) return 0.0;
    SV& s = *static_cast<SV*>(p);
    if (t >= s.n) { g_last_error = 1; return 0.0; }
    return sv_measure_prob(s, t, outcome);
}
// reference core/src/statevector.cpp:190-195
int iq_statevector_measure(void* p, size_t t, double u) {
    if (!p) return 0;
    SV& s = *static_cast<SV*>(p);
    if (t >= s.n) { g_last_error = 1; return 0; }
    const double p0 = sv_measure_prob(s, t, 0);
    const int out = (u < p0) ? 0 : 1;
    if (!sv_collapse(s, t, out)) g_last_error = 2;
    return out;
}
// reference core/src/statevector.cpp:197-204
void iq_statevector_reset_qubit(void* p, size_t t, double u) {
    if (!p) return;
    SV& s = *static_cast<SV*>(p);
    if (t >= s.n) { g_last_error = 1; return; }
    if (iq_statevector_measure(p, t, u) == 1) {
        const cplx X[4] = {0.0, 1.0, 1.0, 0.0};
        sv_apply_1q(s, t, X);
    }
}
// reference core/src/api.cpp:74-83
void iq_statevector_get_amplitudes(void* p, double* re, double* im) {
    if (!p || !re || !im) return;
    const SV& s = *static_cast<SV*>(p);
    for (uint64_t i = 0; i < s.dim; ++i) { re[i] = s.amp[i].real(); im[i] = s.amp[i].imag(); }
}
void iq_statevector_get_probabilities(void* p, double* out) {
    if (!p || !out) return;
    sv_probs(*static_cast<SV*>(p), out);
}
void iq_statevector_sample(void* p, size_t shots, uint64_t seed, uint64_t* out) {
    if (!p || !out) return;
    sv_sample(*static_cast<SV*>(p), shots, seed, out);
}
double iq_statevector_entropy(void* p, const size_t* q, size_t k) {
    if (!p || !q) return 0.0;
    return sv_entropy(*static_cast<SV*>(p), std::vector<size_t>(q, q + k));
}
void iq_statevector_expval_pauli(void* p, const char* ops, const size_t* q, size_t len, double* vr, double* vi) {
    if (!p || !ops || !q || !vr || !vi) return;
    const cplx v = sv_expval_pauli(*static_cast<SV*>(p), ops, q, len);
    *vr = v.real();
    *vi = v.imag();
}
// Reduced density matrix of the kept qubits (reference StateVector::partial_trace,
// core/src/statevector.cpp:293-330, a class method with no C export).
void iqo_statevector_partial_trace(void* p, const size_t* q, size_t k, double* re, double* im) {
    if (!p || !q || !re || !im) return;
    const auto red = sv_partial_trace(*static_cast<SV*>(p), std::vector<size_t>(q, q + k));
    for (size_t i = 0; i < red.size(); ++i) {
        re[i] = red[i].real();
        im[i] = red[i].imag();
    }
}
// Oracle-only helpers (no reference ABI counterpart): load a state, read the norm.
void iqo_statevector_set_amplitudes(void* p, const double* re, const double* im) {
    if (!p || !re || !im) return;
    SV& s = *static_cast<SV*>(p);
    for (uint64_t i = 0; i < s.dim; ++i) s.amp[i] = cplx(re[i], im[i]);
}
double iqo_statevector_norm_squared(void* p) { return p ? sv_norm2(*static_cast<SV*>(p)) : 0.0; }
void iqo_statevector_normalize(void* p) {
    if (p) sv_normalize(*static_cast<SV*>(p));
}
// Raw std::mt19937_64 words and the uniform_real_distribution<double>(0,1)
// draws that reference statevector.cpp:244-250 consumes, for known-answer tests.
void iqo_mt19937_64_raw(uint64_t seed, size_t count, uint64_t* out) {
    std::mt19937_64 gen(seed);
    for (size_t i = 0; i < count; ++i) out[i] = gen();
}
void iqo_uniforms(uint64_t seed, size_t count, double* out) {
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    for (size_t i = 0; i < count; ++i) out[i] = uni(gen);
}

void* iq_density_matrix_create(size_t n) {
    if (n > 16) return nullptr;
    try { return new DM(n); } catch (...) { return nullptr; }
}
void iq_density_matrix_free(void* p) { delete static_cast<DM*>(p); }
void iq_density_matrix_reset(void* p) {
    if (!p) return;
    DM& d = *static_cast<DM*>(p);
    std::fill(d.rho.begin(), d.rho.end(), cplx(0.0, 0.0));
    d.rho[0] = 1.0;
}
void iq_density_matrix_apply_1q_unitary(void* p, size_t t, const double* re, const double* im) {
    if (!p || !re || !im) return;
    DM& d = *static_cast<DM*>(p);
    if (t >= d.n) { g_last_error = 1; return; }
    const auto U = pack(re, im, 4);
    dm_apply_1q(d, t, U.data());
}
void iq_density_matrix_apply_2q_unitary(void* p, size_t q0, size_t q1, const double* re, const double* im) {
    if (!p || !re || !im) return;
    DM& d = *static_cast<DM*>(p);
    if (q0 >= d.n || q1 >= d.n || q0 == q1) { g_last_error = 1; return; }
    const auto U = pack(re, im, 16);
    dm_apply_2q(d, q0, q1, U.data());
}
void iq_density_matrix_apply_kraus_1q(void* p, size_t t, size_t nk, const double* re, const double* im) {
    if (!p || !re || !im) return;
    DM& d = *static_cast<DM*>(p);
    if (t >= d.n) { g_last_error = 1; return; }
    dm_kraus_1q(d, t, pack(re, im, 4 * nk));
}
// reference core/src/api.cpp:157-167
void iq_density_matrix_get_elements(void* p, double* re, double* im) {
    if (!p || !re || !im) return;
    const DM& d = *static_cast<DM*>(p);
    const uint64_t total = d.dim * d.dim;
    for (uint64_t i = 0; i < total; ++i) { re[i] = d.rho[i].real(); im[i] = d.rho[i].imag(); }
}
double iq_density_matrix_trace(void* p) { return p ? dm_trace(*static_cast<DM*>(p)) : 0.0; }
double iq_density_matrix_purity(void* p) { return p ? dm_purity(*static_cast<DM*>(p)) : 0.0; }
double iq_density_matrix_entropy(void* p) { return p ? dm_entropy(*static_cast<DM*>(p)) : 0.0; }
int iq_density_matrix_measure(void* p, size_t t, double u) {
    if (!p) return 0;
    DM& d = *static_cast<DM*>(p);
    if (t >= d.n) { g_last_error = 1; return 0; }
    const double p0 = dm_measure_prob(d, t, 0);
    const int out = (u < p0) ? 0 : 1;
    if (!dm_collapse(d, t, out)) g_last_error = 2;
    return out;
}
// reference core/src/density_matrix.cpp:258-269 — Kraus {|0><0|, |0><1|}.
void iq_density_matrix_reset_qubit(void* p, size_t t) {
    if (!p) return;
    DM& d = *static_cast<DM*>(p);
    if (t >= d.n) { g_last_error = 1; return; }
    const std::vector<cplx> ks = {1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
    dm_kraus_1q(d, t, ks);
}
// Oracle-only: <Z_i> for every qubit from the diagonal, as the reference's
// caller reads it (invariantq/simulator/density_matrix.py:60-69 takes diag(rho)).
void iqo_density_matrix_expval_z(void* p, double* out) {
    if (!p || !out) return;
    const DM& d = *static_cast<DM*>(p);
    for (size_t q = 0; q < d.n; ++q) out[q] = 0.0;
    for (uint64_t i = 0; i < d.dim; ++i) {
        const double v = d.rho[i * d.dim + i].real();
        for (size_t q = 0; q < d.n; ++q) out[q] += ((i >> q) & 1ULL) ? -v : v;
    }
}
void iqo_density_matrix_set_elements(void* p, const double* re, const double* im) {
    if (!p || !re || !im) return;
    DM& d = *static_cast<DM*>(p);
    const uint64_t total = d.dim * d.dim;
    for (uint64_t i = 0; i < total; ++i) d.rho[i] = cplx(re[i], im[i]);
}

}  // extern "C"
