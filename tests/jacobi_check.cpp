// CPU check of invariantq_amd/csrc/jacobi.hpp (row-maximum pivot search)
// against the oracle's full-scan restatement of the reference's capped
// Jacobi (core/src/statevector.cpp:333-393), bit for bit, on random Hermitian
// matrices, rank-deficient density matrices and matrices with exactly tied
// off-diagonal moduli.  Prints "ok" or the first mismatch.
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
            cases += 4;
        }
    }
    if (bad) {
        printf("%d of %d cases differ\n", bad, cases);
        return 1;
    }
    printf("ok %d cases\n", cases);
    return 0;
}
