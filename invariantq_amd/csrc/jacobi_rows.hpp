// The capped largest-pivot Jacobi of jacobi.hpp with the matrix held
// elsewhere (the device): the same pivots, rotations and eigenvalues, bit for
// bit, where the host never holds the d x d matrix.
//
// The host keeps, per row i, the largest modulus right of the diagonal
// (rmax), the column of its first occurrence (rarg), that entry's squared
// modulus (rsq) and its value (rval), plus the real diagonal.  The pivot
// choice, the rotation parameters and every std::abs are host work, as in
// jacobi.hpp.  The backend does what touches O(d) or O(d^2) entries:
//   rotate:  apply the rotation (p, q) to the matrix and return the new
//            entries (i, p) and (i, q) for i < q;
//   scan:    for each listed row, the entries right of the diagonal that can
//            hold the row's largest std::abs: with top the largest squared
//            modulus, those whose squared modulus is at least
//            top * kJacobiSlack (every non-zero entry when top is below
//            kJacobiFloor), at most kRowCands of them per row, or, when there
//            are more and they all have the same unordered pair
//            (|re|, |im|), so the same std::abs, the first of them;
//   fetch:   one row right of its diagonal, for the host to scan itself when
//            neither applies.
// A candidate list is complete for the row's largest std::abs and its first
// column (an entry below the filter bound has a smaller std::abs than the
// row's largest squared modulus allows, as in jacobi.hpp), so the host's
// choice among the candidates in column order is the full scan's.
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <vector>

#include "jacobi.hpp"

namespace iqg {

constexpr uint32_t kRowCands = 16;        // candidates returned per row
constexpr uint32_t kRowFetch = 0xffffffffu;  // scan count: the host reads the row

// Results of scanning `nrows` listed rows: count[t] candidates for rows[t]
// (kRowFetch: read the row), the first in first_j / first_v, the others in
// the shared list extra_t / extra_j / extra_v (extra_n entries, of which the
// first cap were stored; row order within the list is arbitrary).
struct RowScan {
    const uint32_t* count = nullptr;
    const uint32_t* first_j = nullptr;
    const std::complex<double>* first_v = nullptr;
    uint32_t extra_n = 0, extra_cap = 0;
    const uint32_t* extra_t = nullptr;
    const uint32_t* extra_j = nullptr;
    const std::complex<double>* extra_v = nullptr;
};

// Backend B:
//   bool diagonal(double* re)                          real parts of A(i, i)
//   bool scan(const uint32_t* rows, size_t nrows, RowScan& out)
//   bool rotate_scan(size_t p, size_t q, double c, std::complex<double> sn, double npp, double nqq,
//                    const std::complex<double>*& colp, const std::complex<double>*& colq,
//                    const uint32_t* rows, size_t nrows, RowScan& out)
//       rotation as jacobi.hpp's loop (A(p, p) = npp, A(q, q) = nqq, A(p, q) = A(q, p) = 0),
//       colp[i] = new A(i, p), colq[i] = new A(i, q) for i < q, then scan as above
//   bool fetch_row(size_t i, std::complex<double>* out)  out[j - i - 1] = A(i, j), j > i
// Returns false (and an empty vector) when the backend fails.
template <class B>
bool jacobi_eigenvalues_rows(B& dev, size_t d, int max_iter, double tol, std::vector<double>& ev) {
    using cplx = std::complex<double>;
    constexpr double kSlack = kJacobiSlack, kFloor = kJacobiFloor;
    auto sq = [](const cplx& z) { return z.real() * z.real() + z.imag() * z.imag(); };
    ev.clear();
    std::vector<double> rmax(d, 0.0), rsq(d, 0.0), diag(d, 0.0);
    std::vector<size_t> rarg(d, 0);
    std::vector<cplx> rval(d, cplx(0.0, 0.0)), row;
    if (d < 2) {
        if (d == 1 && !dev.diagonal(diag.data())) return false;
        ev = diag;
        return true;
    }
    if (!dev.diagonal(diag.data())) return false;

    // the row's entry of largest std::abs among candidates (first column among
    // equals), as jacobi.hpp's rescan
    auto settle = [&](size_t i, const std::vector<std::pair<uint32_t, cplx>>& cands) {
        double best = 0.0;
        size_t arg = i + 1;
        cplx val(0.0, 0.0);
        for (const auto& c : cands) {
            const double v = std::abs(c.second);
            if (v > best) {
                best = v;
                arg = c.first;
                val = c.second;
            }
        }
        rmax[i] = best;
        rarg[i] = arg;
        rval[i] = val;
        rsq[i] = sq(val);
    };
    // the full scan of a fetched row (jacobi.hpp's rescan on it)
    auto scan_fetched = [&](size_t i) -> bool {
        row.resize(d - i - 1);
        if (!dev.fetch_row(i, row.data())) return false;
        double top = 0.0;
        for (const cplx& z : row) top = std::max(top, sq(z));
        const double bound = top >= kFloor ? top * kSlack : -1.0;
        std::vector<std::pair<uint32_t, cplx>> cands;
        for (size_t j = i + 1; j < d; ++j)
            if (sq(row[j - i - 1]) >= bound) cands.emplace_back((uint32_t)j, row[j - i - 1]);
        settle(i, cands);
        return true;
    };
    std::vector<std::vector<std::pair<uint32_t, cplx>>> extra;
    auto apply_scan = [&](const uint32_t* rows, size_t nrows, const RowScan& rs) -> bool {
        const bool lost = rs.extra_n > rs.extra_cap;  // the shared list overflowed
        extra.assign(nrows, {});
        for (uint32_t k = 0; k < std::min(rs.extra_n, rs.extra_cap); ++k)
            extra[rs.extra_t[k]].emplace_back(rs.extra_j[k], rs.extra_v[k]);
        std::vector<std::pair<uint32_t, cplx>> cands;
        for (size_t t = 0; t < nrows; ++t) {
            const size_t i = rows[t];
            const uint32_t n = rs.count[t];
            if (n == kRowFetch || (lost && n > 1) || (n > 1 && extra[t].size() != n - 1)) {
                if (!scan_fetched(i)) return false;
                continue;
            }
            cands.clear();
            if (n >= 1) cands.emplace_back(rs.first_j[t], rs.first_v[t]);
            cands.insert(cands.end(), extra[t].begin(), extra[t].end());
            std::sort(cands.begin(), cands.end(),
                      [](const std::pair<uint32_t, cplx>& a, const std::pair<uint32_t, cplx>& b) { return a.first < b.first; });
            settle(i, cands);
        }
        return true;
    };

    std::vector<uint32_t> rows;
    rows.reserve(d);
    for (size_t i = 0; i + 1 < d; ++i) rows.push_back((uint32_t)i);
    {
        RowScan rs;
        if (!dev.scan(rows.data(), rows.size(), rs) || !apply_scan(rows.data(), rows.size(), rs)) return false;
    }
    std::vector<char> listed(d, 0);
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
        // rval[p] is A(p, q): the row's recorded maximum, rescanned whenever
        // its column or its row changed since
        jacobi_rotation(diag[p], diag[q], rval[p], c, sn, npp, nqq);
        // rows whose recorded maximum sits in a changed column are rescanned,
        // with rows p and q; the others are offered their new entries
        rows.clear();
        for (size_t i = 0; i < q; ++i) {
            listed[i] = 0;
            if (i == p) continue;
            if (rarg[i] == q || (i < p && rarg[i] == p)) {
                rows.push_back((uint32_t)i);
                listed[i] = 1;
            }
        }
        rows.push_back((uint32_t)p);
        if (q + 1 < d) rows.push_back((uint32_t)q);
        const cplx *colp = nullptr, *colq = nullptr;
        RowScan rs;
        if (!dev.rotate_scan(p, q, c, sn, npp, nqq, colp, colq, rows.data(), rows.size(), rs)) return false;
        diag[p] = npp;
        diag[q] = nqq;
        for (size_t i = 0; i < q; ++i) {
            if (i == p || listed[i]) continue;
            auto offer = [&](size_t j, const cplx& z) {
                const double s2 = sq(z);
                if (rsq[i] >= kFloor && s2 < rsq[i] * kSlack) return;  // below the recorded maximum
                const double v = std::abs(z);
                if (v > rmax[i] || (v == rmax[i] && v > 0.0 && j < rarg[i])) {
                    rmax[i] = v;
                    rarg[i] = j;
                    rsq[i] = s2;
                    rval[i] = z;
                }
            };
            if (i < p) offer(p, colp[i]);
            offer(q, colq[i]);
        }
        if (!apply_scan(rows.data(), rows.size(), rs)) return false;
    }
    ev = diag;
    return true;
}

}  // namespace iqg
