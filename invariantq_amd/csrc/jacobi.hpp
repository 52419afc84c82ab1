// Eigenvalues of a Hermitian matrix by the capped largest-pivot complex Jacobi
// method of reference core/src/statevector.cpp:333-393 (and
// core/src/density_matrix.cpp:322-370), kept so entropies agree with it:
// each rotation zeroes the off-diagonal entry of largest modulus (the first
// in row-major order among equals), at most `max_iter` rotations, stopping
// once that modulus is below `tol`.
//
// The reference rescans the whole upper triangle for every pivot, O(d^2) per
// rotation (seconds at d = 4096).  Here each row keeps the largest modulus
// right of its diagonal and the column of its first occurrence.  A rotation at
// (p, q) changes rows p and q and the entries (i, p), (i, q) of the rows
// above them, so only those are updated, and a row is rescanned only when its
// recorded maximum sat in a changed column.
//
// The moduli are std::abs values (a hypot call, about twenty times the cost
// of a squared modulus).  hypot is within 1 ulp of |z| and re^2 + im^2 within
// a few ulps of |z|^2, so an entry whose squared modulus is below
// (1 - 1e-14) times the row's largest cannot have the row's largest (or a
// tied) std::abs: a scan computes std::abs only for the entries above that
// bound.  Rows whose largest squared modulus is below 1e-280, where squares
// lose precision to underflow, are scanned with std::abs throughout.  The
// pivots, and therefore the arithmetic and the result, are those of the full
// scan, bit for bit.
//
// The first scan of every row (2^29 entries for a 15-qubit reduced matrix)
// and each rotation's row updates are split over host threads, rows dealt
// round-robin so each thread gets a similar share of the triangle; rows are
// independent, so the result does not depend on the split.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <thread>
#include <vector>

namespace iqg {

constexpr double kJacobiSlack = 1.0 - 1e-14;  // squared-modulus filter, relative
constexpr double kJacobiFloor = 1e-280;       // squared moduli below this: no filter

// One rotation's parameters for pivot (p, q) of the capped Jacobi: the
// cosine c and the complex sine sn that zero A(p, q), and the new diagonal
// entries, from app = A(p, p), aqq = A(q, q) (real parts) and apq = A(p, q).
inline void jacobi_rotation(double app, double aqq, std::complex<double> apq, double& c, std::complex<double>& sn,
                            double& npp, double& nqq) {
    using cplx = std::complex<double>;
    const double th = 0.5 * std::atan2(2.0 * std::abs(apq), app - aqq);
    const cplx ph = apq / std::abs(apq);
    c = std::cos(th);
    sn = std::sin(th) * ph;
    const double cross = 2.0 * (c * std::conj(sn) * apq).real();
    npp = c * c * app + std::norm(sn) * aqq + cross;
    nqq = std::norm(sn) * app + c * c * aqq - cross;
}

// Eigenvalue estimates of the d x d Hermitian matrix A (row-major), which is
// overwritten.
inline std::vector<double> jacobi_eigenvalues_inplace(std::complex<double>* A, size_t d, int max_iter, double tol) {
    using cplx = std::complex<double>;
    auto at = [&](size_t i, size_t j) -> cplx& { return A[i * d + j]; };
    constexpr double kSlack = kJacobiSlack, kFloor = kJacobiFloor;
    auto sq = [](const cplx& z) { return z.real() * z.real() + z.imag() * z.imag(); };
    std::vector<double> rmax(d, 0.0);  // largest |A(i, j)|, j > i
    std::vector<size_t> rarg(d, 0);    // first j attaining it
    std::vector<double> rsq(d, 0.0);   // squared modulus of A(i, rarg[i])
    auto rescan = [&](size_t i) {
        double top = 0.0;
        for (size_t j = i + 1; j < d; ++j) top = std::max(top, sq(at(i, j)));
        const double bound = top >= kFloor ? top * kSlack : -1.0;
        double best = 0.0;
        size_t arg = i + 1;
        for (size_t j = i + 1; j < d; ++j) {
            if (sq(at(i, j)) < bound) continue;
            const double v = std::abs(at(i, j));
            if (v > best) {
                best = v;
                arg = j;
            }
        }
        rmax[i] = best;
        rarg[i] = arg;
        rsq[i] = arg < d ? sq(at(i, arg)) : 0.0;
    };
    // row i's entries in columns j1 (and j2) changed: rescan when its recorded
    // maximum sat in one of them, else offer the new values (ties go to the
    // smaller column, as in a row-major scan)
    auto update = [&](size_t i, size_t j1, size_t j2, bool two) {
        if (rarg[i] == j1 || (two && rarg[i] == j2)) {
            rescan(i);
            return;
        }
        auto offer = [&](size_t j) {
            const double s2 = sq(at(i, j));
            if (rsq[i] >= kFloor && s2 < rsq[i] * kSlack) return;  // below the recorded maximum
            const double v = std::abs(at(i, j));
            if (v > rmax[i] || (v == rmax[i] && v > 0.0 && j < rarg[i])) {
                rmax[i] = v;
                rarg[i] = j;
                rsq[i] = s2;
            }
        };
        offer(j1);
        if (two) offer(j2);
    };
    // rows dealt round-robin over host threads (rows are independent: each
    // writes only its own row maximum), on the calling thread for small d or
    // when no thread can be had
    const size_t hw = std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
    auto for_rows = [&](size_t end, size_t nthreads, auto&& body) {
        auto part = [&](size_t t) {
            for (size_t i = t; i < end; i += nthreads) body(i);
        };
        std::vector<std::thread> pool;
        size_t t = 1;
        try {
            for (; t < nthreads; ++t) pool.emplace_back(part, t);
        } catch (const std::exception&) {
            for (; t < nthreads; ++t) part(t);
        }
        part(0);
        for (auto& th : pool) th.join();
    };
    for_rows(d - 1, d >= 512 ? hw : 1, [&](size_t i) { rescan(i); });
    // a rotation's updates cost about 20 us per thread started: threads from
    // d = 2048 (below, the whole 100 rotations take a few ms on one thread)
    const size_t rot_threads = d >= 2048 ? hw : 1;
    for (int it = 0; it < max_iter; ++it) {
        double best = 0.0;
        size_t p = 0, q = 1;
        for (size_t i = 0; i + 1 < d; ++i)
            if (rmax[i] > best) {
                best = rmax[i];
                p = i;
                q = rarg[i];
            }
        if (best < tol) break;
        double c, npp, nqq;
        cplx sn;
        jacobi_rotation(at(p, p).real(), at(q, q).real(), at(p, q), c, sn, npp, nqq);
        for (size_t i = 0; i < d; ++i) {
            if (i == p || i == q) continue;
            const cplx aip = at(i, p), aiq = at(i, q);
            at(i, p) = c * aip + std::conj(sn) * aiq;
            at(p, i) = std::conj(at(i, p));
            at(i, q) = -sn * aip + c * aiq;
            at(q, i) = std::conj(at(i, q));
        }
        at(p, p) = npp;
        at(q, q) = nqq;
        at(p, q) = 0.0;
        at(q, p) = 0.0;
        // rows above p and q: their entries in columns p and q changed (a row
        // whose maximum sat there is rescanned, O(d): at d = 2^15 about 13,600
        // rows per rotation, so the rows go over host threads)
        for_rows(q, rot_threads, [&](size_t i) {
            if (i == p) return;
            if (i < p) update(i, p, q, true);
            else update(i, q, q, false);
        });
        rescan(p);
        if (q + 1 < d) rescan(q);
    }
    std::vector<double> ev(d);
    for (size_t i = 0; i < d; ++i) ev[i] = at(i, i).real();
    return ev;
}

inline std::vector<double> jacobi_eigenvalues(std::vector<std::complex<double>> A, size_t d, int max_iter,
                                              double tol) {
    return jacobi_eigenvalues_inplace(A.data(), d, max_iter, tol);
}

inline double entropy_bits(const std::vector<double>& ev) {
    double h = 0.0;
    for (double l : ev)
        if (l > 1e-14) h -= l * std::log2(l);
    return h > 0.0 ? h : 0.0;
}

}  // namespace iqg
