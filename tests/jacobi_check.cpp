// CPU check of invariantq_amd/csrc/jacobi.hpp (row-maximum pivot search)
// against the oracle's full-scan restatement of the reference's capped
// Jacobi (core/src/statevector.cpp:333-393), bit for bit, on random Hermitian
// matrices, rank-deficient density matrices, matrices with exactly tied and
// 1-ulp-apart off-diagonal moduli, matrices whose squared moduli underflow,
// pairs with equal std::abs but differently ordered squares (the edge of the
// squared-modulus filter), and a 1024 x 1024 half-cut reduced state.  The same
// cases go through jacobi_rows.hpp's driver (the device form: the host holds
// per-row records, a backend rotates and scans rows) with a host backend that
// does what the device kernels do: candidate lists capped per row, one shared
// list of further candidates (with a small cap, so that it overflows), rows
// of one repeated modulus pair reduced to their first entry, fetched rows.
// Prints "ok" or the first mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "jacobi.hpp"
#include "jacobi_rows.hpp"

// Host stand-in for the device's rotate and scan kernels (the semantics of
// k_jacobi_rotate / k_jacobi_scan), on a private copy of the matrix.
struct HostRows {
    using cplx = std::complex<double>;
    std::vector<cplx> A;
    size_t d;
    uint32_t cap;
    std::vector<uint32_t> cnt, fj, et, ej;
    std::vector<cplx> fv, ev_, colp, colq;
    size_t fetches = 0, uniform = 0, overflowed = 0;
    HostRows(const std::vector<cplx>& M, size_t d_, uint32_t cap_) : A(M), d(d_), cap(cap_) {}
    cplx& at(size_t i, size_t j) { return A[i * d + j]; }
    bool diagonal(double* re) {
        for (size_t i = 0; i < d; ++i) re[i] = at(i, i).real();
        return true;
    }
    bool scan(const uint32_t* rows, size_t nrows, iqg::RowScan& out) {
        cnt.assign(nrows, 0);
        fj.assign(nrows, 0);
        fv.assign(nrows, 0.0);
        et.clear();
        ej.clear();
        ev_.clear();
        uint32_t en = 0;
        for (size_t t = 0; t < nrows; ++t) {
            const size_t i = rows[t];
            double top = 0.0;
            for (size_t j = i + 1; j < d; ++j) {
                const cplx z = at(i, j);
                top = std::max(top, z.real() * z.real() + z.imag() * z.imag());
            }
            const double bound = top >= iqg::kJacobiFloor ? top * iqg::kJacobiSlack : -1.0;
            std::vector<size_t> c;
            for (size_t j = i + 1; j < d; ++j) {
                const cplx z = at(i, j);
                const double s2 = z.real() * z.real() + z.imag() * z.imag();
                if (s2 >= bound && (z.real() != 0.0 || z.imag() != 0.0)) c.push_back(j);
            }
            // the device's candidates arrive in no particular order
            std::reverse(c.begin(), c.end());
            if (c.size() > iqg::kRowCands) {
                size_t first = d;
                for (size_t j : c) first = std::min(first, j);
                auto key = [&](size_t j) {
                    const double a = std::fabs(at(i, j).real()), b = std::fabs(at(i, j).imag());
                    return std::make_pair(std::max(a, b), std::min(a, b));
                };
                bool same = true;
                for (size_t j : c) same &= key(j) == key(first);
                if (same) {
                    cnt[t] = 1;
                    fj[t] = (uint32_t)first;
                    fv[t] = at(i, first);
                    ++uniform;
                } else {
                    cnt[t] = iqg::kRowFetch;
                }
                continue;
            }
            cnt[t] = (uint32_t)c.size();
            for (size_t k = 0; k < c.size(); ++k) {
                if (k == 0) {
                    fj[t] = (uint32_t)c[k];
                    fv[t] = at(i, c[k]);
                } else {
                    if (en < cap) {
                        et.push_back((uint32_t)t);
                        ej.push_back((uint32_t)c[k]);
                        ev_.push_back(at(i, c[k]));
                    }
                    ++en;
                }
            }
        }
        out.count = cnt.data();
        out.first_j = fj.data();
        out.first_v = fv.data();
        if (en > cap) ++overflowed;
        out.extra_n = en;
        out.extra_cap = cap;
        out.extra_t = et.data();
        out.extra_j = ej.data();
        out.extra_v = ev_.data();
        return true;
    }
    bool rotate_scan(size_t p, size_t q, double c, cplx sn, double npp, double nqq, const cplx*& cp, const cplx*& cq,
                     const uint32_t* rows, size_t nrows, iqg::RowScan& out) {
        colp.assign(d, 0.0);
        colq.assign(d, 0.0);
        for (size_t i = 0; i < d; ++i) {
            if (i == p || i == q) continue;
            const cplx aip = at(i, p), aiq = at(i, q);
            at(i, p) = c * aip + std::conj(sn) * aiq;
            at(p, i) = std::conj(at(i, p));
            at(i, q) = -sn * aip + c * aiq;
            at(q, i) = std::conj(at(i, q));
            colp[i] = at(i, p);
            colq[i] = at(i, q);
        }
        at(p, p) = npp;
        at(q, q) = nqq;
        at(p, q) = 0.0;
        at(q, p) = 0.0;
        cp = colp.data();
        cq = colq.data();
        return scan(rows, nrows, out);
    }
    bool fetch_row(size_t i, cplx* out) {
        ++fetches;
        for (size_t j = i + 1; j < d; ++j) out[j - i - 1] = at(i, j);
        return true;
    }
};

extern "C" void iqo_hermitian_eigenvalues(const double* re, const double* im, size_t d, int max_iter, double tol,
                                          double* out);

static size_t g_fetched = 0, g_uniform = 0, g_overflowed = 0;

static int check(const std::vector<std::complex<double>>& A, size_t d, int max_iter, double tol, const char* what) {
    std::vector<double> re(d * d), im(d * d), ref(d);
    for (size_t i = 0; i < d * d; ++i) {
        re[i] = A[i].real();
        im[i] = A[i].imag();
    }
    iqo_hermitian_eigenvalues(re.data(), im.data(), d, max_iter, tol, ref.data());
    const auto got = iqg::jacobi_eigenvalues(A, d, max_iter, tol);
    auto same = [&](const std::vector<double>& v, const char* form) {
        if (v.size() == d && std::memcmp(v.data(), ref.data(), d * sizeof(double)) == 0) return 0;
        for (size_t i = 0; i < d; ++i)
            if (i >= v.size() || v[i] != ref[i]) {
                printf("mismatch %s (%s) d=%zu iters=%d at %zu: %.17g vs %.17g\n", what, form, d, max_iter, i,
                       i < v.size() ? v[i] : NAN, ref[i]);
                break;
            }
        return 1;
    };
    int bad = same(got, "row maxima");
    // the device form, with a shared candidate list of 40 (overflows on tied
    // matrices) and of 1 << 20
    for (uint32_t cap : {40u, 1u << 20}) {
        HostRows dev(A, d, cap);
        std::vector<double> rows_ev;
        if (!iqg::jacobi_eigenvalues_rows(dev, d, max_iter, tol, rows_ev)) {
            printf("device form failed %s d=%zu\n", what, d);
            return 1;
        }
        bad |= same(rows_ev, cap == 40u ? "rows, cap 40" : "rows");
        g_fetched += dev.fetches;
        g_uniform += dev.uniform;
        g_overflowed += dev.overflowed;
    }
    return bad;
}

int main() {
    std::mt19937_64 g(12345);
    std::normal_distribution<double> N(0.0, 1.0);
    int bad = 0, cases = 0;
    for (size_t d : {2, 3, 4, 7, 16, 33, 64, 128, 256}) {
        for (int iters : {100, 120, 5000}) {
            // random Hermitian
            std::vector<std::complex<double>> A(d * d);
            for (size_t i = 0; i < d; ++i)
                for (size_t j = i; j < d; ++j) {
                    const std::complex<double> v(N(g), i == j ? 0.0 : N(g));
                    A[i * d + j] = v;
                    A[j * d + i] = std::conj(v);
                }
            bad += check(A, d, iters, 1e-12, "hermitian");
            // rank-3 density matrix (a reduced state)
            std::vector<std::complex<double>> R(d * d, 0.0);
            for (int r = 0; r < 3; ++r) {
                std::vector<std::complex<double>> v(d);
                double nrm = 0.0;
                for (auto& x : v) {
                    x = {N(g), N(g)};
                    nrm += std::norm(x);
                }
                for (size_t i = 0; i < d; ++i)
                    for (size_t j = 0; j < d; ++j) R[i * d + j] += v[i] * std::conj(v[j]) / (3.0 * nrm);
            }
            bad += check(R, d, iters, 1e-12, "density");
            // exact ties: entries from a small set of moduli with signs and i
            std::uniform_int_distribution<int> U(0, 3);
            std::vector<std::complex<double>> T(d * d, 0.0);
            for (size_t i = 0; i < d; ++i)
                for (size_t j = i; j < d; ++j) {
                    const double mag = 0.25 * (1 + U(g) % 2);
                    const int ph = U(g);
                    std::complex<double> v = ph == 0 ? std::complex<double>(mag, 0) : ph == 1 ? std::complex<double>(-mag, 0)
                                             : ph == 2 ? std::complex<double>(0, mag) : std::complex<double>(0, -mag);
                    if (i == j) v = {0.5 * U(g), 0.0};
                    T[i * d + j] = v;
                    T[j * d + i] = std::conj(v);
                }
            bad += check(T, d, iters, 1e-13, "ties");
            // GHZ-like: two equal blocks
            std::vector<std::complex<double>> Gm(d * d, 0.0);
            Gm[0] = Gm[d * d - 1] = 0.5;
            Gm[d - 1] = Gm[(d - 1) * d] = 0.5;
            bad += check(Gm, d, iters, 1e-12, "ghz");
            // near ties: moduli 1 ulp apart, with phases that make hypot round
            std::vector<std::complex<double>> Nt(d * d, 0.0);
            const double m0 = 0.25, mags[3] = {std::nextafter(m0, 0.0), m0, std::nextafter(m0, 1.0)};
            for (size_t i = 0; i < d; ++i)
                for (size_t j = i; j < d; ++j) {
                    const double mag = mags[U(g) % 3];
                    const double c = std::cos(0.1 * U(g) + 0.05), s = std::sin(0.1 * U(g) + 0.05);
                    std::complex<double> v = U(g) < 2 ? std::complex<double>(mag, 0.0) : std::complex<double>(mag * c, mag * s);
                    if (i == j) v = {0.5 * U(g), 0.0};
                    Nt[i * d + j] = v;
                    Nt[j * d + i] = std::conj(v);
                }
            bad += check(Nt, d, iters, 1e-13, "near-ties");
            // squared moduli below the underflow floor: a scaled random matrix
            // (every row) and one with only some rows scaled
            std::vector<std::complex<double>> S(A), M(A);
            for (size_t i = 0; i < d; ++i)
                for (size_t j = 0; j < d; ++j) {
                    S[i * d + j] *= 1e-160;
                    if ((i | j) & 1) M[i * d + j] *= 1e-150;
                }
            bad += check(S, d, iters, 1e-300, "tiny");
            bad += check(M, d, iters, 1e-300, "mixed-tiny");
            cases += 7;
        }
    }
    {
        // the squared-modulus filter's edge: two entries of row 0 with equal
        // std::abs, the later one with the larger re^2 + im^2 (in both the
        // plain and the fused evaluation).  The full scan takes the earlier.
        std::uniform_real_distribution<double> ang(0.1, 1.4), mod(0.29, 0.31);
        int found = 0;
        for (int attempt = 0; attempt < 100000 && found < 8; ++attempt) {
            const double r = mod(g), t = ang(g);
            const std::complex<double> a(r * std::cos(t), r * std::sin(t));
            const double m = std::abs(a);
            const std::complex<double> b(m, 0.0);
            volatile double x2 = a.real() * a.real(), y2 = a.imag() * a.imag(), m2 = m * m;
            const double sa_plain = x2 + y2, sa_fma = std::fma(a.real(), a.real(), y2);
            if (std::abs(b) != m || !(sa_plain < m2) || !(sa_fma < m2)) continue;
            for (size_t d : {3, 16}) {
                std::vector<std::complex<double>> E(d * d, 0.0);
                for (size_t i = 0; i < d; ++i)
                    for (size_t j = i; j < d; ++j) {
                        std::complex<double> v(0.01 * N(g), i == j ? 0.0 : 0.01 * N(g));
                        if (i == 0 && j == 1) v = a;
                        if (i == 0 && j == 2) v = b;
                        E[i * d + j] = v;
                        E[j * d + i] = std::conj(v);
                    }
                bad += check(E, d, 100, 1e-12, "filter-edge");
                ++cases;
            }
            ++found;
        }
        if (found < 8) {
            printf("filter-edge: only %d pairs found\n", found);
            return 1;
        }
    }
    {
        // a 10-qubit half cut of a random 20-qubit state (d = 1024, the
        // reference's 100-rotation cap)
        const size_t d = 1024, e = 1024;
        std::vector<std::complex<double>> psi(d * e), R(d * d);
        double nrm = 0.0;
        for (auto& x : psi) {
            x = {N(g), N(g)};
            nrm += std::norm(x);
        }
        for (auto& x : psi) x /= std::sqrt(nrm);
        for (size_t i = 0; i < d; ++i)
            for (size_t j = 0; j < d; ++j) {
                std::complex<double> a = 0.0;
                for (size_t r = 0; r < e; ++r) a += psi[i * e + r] * std::conj(psi[j * e + r]);
                R[i * d + j] = a;
            }
        bad += check(R, d, 100, 1e-12, "half-cut-1024");
        ++cases;
    }
    {
        // rows scanned by several host threads (d >= 512): exact ties over a
        // 600 x 600 matrix, and a rank-8 reduced state at d = 2048 (an 11-qubit
        // cut), against the single-threaded full scan
        std::uniform_int_distribution<int> U(0, 3);
        const size_t d = 600;
        std::vector<std::complex<double>> T(d * d, 0.0);
        for (size_t i = 0; i < d; ++i)
            for (size_t j = i; j < d; ++j) {
                const double mag = 0.25 * (1 + U(g) % 2);
                std::complex<double> v = (U(g) & 1) ? std::complex<double>(0, mag) : std::complex<double>(-mag, 0);
                if (i == j) v = {0.5 * U(g), 0.0};
                T[i * d + j] = v;
                T[j * d + i] = std::conj(v);
            }
        bad += check(T, d, 100, 1e-13, "ties-600");
        const size_t d2 = 2048, r = 8;
        std::vector<std::complex<double>> V(d2 * r), R(d2 * d2);
        for (auto& x : V) x = {N(g), N(g)};
        for (size_t i = 0; i < d2; ++i)
            for (size_t j = 0; j < d2; ++j) {
                std::complex<double> a = 0.0;
                for (size_t c = 0; c < r; ++c) a += V[i * r + c] * std::conj(V[j * r + c]);
                R[i * d2 + j] = a * (1.0 / (d2 * r));
            }
        bad += check(R, d2, 100, 1e-12, "rank8-2048");
        cases += 2;
    }
    if (bad) {
        printf("%d of %d cases differ\n", bad, cases);
        return 1;
    }
    // each path of the device form was taken
    if (!g_fetched || !g_uniform || !g_overflowed) {
        printf("device form: fetched %zu, uniform %zu, overflowed %zu: a path was not exercised\n", g_fetched, g_uniform,
               g_overflowed);
        return 1;
    }
    printf("ok %d cases (device form: %zu rows fetched, %zu tied rows reduced, %zu overflowed lists)\n", cases,
           g_fetched, g_uniform, g_overflowed);
    return 0;
}
