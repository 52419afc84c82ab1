// CPU check of invariantq_amd/csrc/jacobi.hpp (row-maximum pivot search)
// against the oracle's full-scan restatement of the reference's capped
// Jacobi (core/src/statevector.cpp:333-393), bit for bit, on random Hermitian
// matrices, rank-deficient density matrices, matrices with exactly tied and
// 1-ulp-apart off-diagonal moduli, matrices whose squared moduli underflow,
// pairs with equal std::abs but differently ordered squares (the edge of the
// squared-modulus filter), and a 1024 x 1024 half-cut reduced state.  Prints
// "ok" or the first mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "jacobi.hpp"

extern "C" void iqo_hermitian_eigenvalues(const double* re, const double* im, size_t d, int max_iter, double tol,
                                          double* out);

static int check(const std::vector<std::complex<double>>& A, size_t d, int max_iter, double tol, const char* what) {
    std::vector<double> re(d * d), im(d * d), ref(d);
    for (size_t i = 0; i < d * d; ++i) {
        re[i] = A[i].real();
        im[i] = A[i].imag();
    }
    iqo_hermitian_eigenvalues(re.data(), im.data(), d, max_iter, tol, ref.data());
    const auto got = iqg::jacobi_eigenvalues(A, d, max_iter, tol);
    if (std::memcmp(got.data(), ref.data(), d * sizeof(double)) != 0) {
        for (size_t i = 0; i < d; ++i)
            if (got[i] != ref[i]) {
                printf("mismatch %s d=%zu iters=%d at %zu: %.17g vs %.17g\n", what, d, max_iter, i, got[i], ref[i]);
                break;
            }
        return 1;
    }
    return 0;
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
    printf("ok %d cases\n", cases);
    return 0;
}
