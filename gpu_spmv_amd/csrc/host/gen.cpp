// Synthetic benchmark matrices.  The reference reads SuiteSparse files only;
// these generators produce the seeded inputs the benchmark configurations
// describe.  Every row draws from its own RNG stream (row i -> stream i+1),
// so the output is independent of the OpenMP thread count.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "common.hpp"
#include "rng.hpp"

using namespace axspmv;

namespace {

// Finishes a matrix whose row[] holds row lengths at row[1..n]: prefix sum.
int finish_rows(axspmv_csr *a) {
    for (uint32_t i = 0; i < a->nrows; ++i) a->row[i + 1] += a->row[i];
    return AXSPMV_OK;
}

// Allocates a CSR from per-row lengths (64-bit total checked against 2^32).
int alloc_from_lengths(uint32_t nrows, uint32_t ncols, const std::vector<uint32_t> &rl, axspmv_csr **out) {
    uint64_t total = 0;
    for (uint32_t v : rl) total += v;
    if (total >= (1ull << 32)) return fail(AXSPMV_E_OVERFLOW, "generator: nnz %llu exceeds 32-bit indices", (unsigned long long)total);
    AX_TRY(axspmv_csr_alloc(nrows, ncols, uint32_t(total), out));
    for (uint32_t i = 0; i < nrows; ++i) (*out)->row[i + 1] = rl[i];
    return finish_rows(*out);
}

// `k` distinct uniform draws from [0, n), sorted, appended to `cols`.
void distinct_sorted(Rng &r, uint32_t n, uint32_t k, std::vector<uint32_t> &cols) {
    cols.clear();
    if (k == 0) return;
    if (uint64_t(k) * 4 >= n) {
        // Dense case: partial Fisher-Yates over [0, n).
        std::vector<uint32_t> all(n);
        std::iota(all.begin(), all.end(), 0u);
        for (uint32_t i = 0; i < k; ++i) {
            const uint32_t j = i + uint32_t(r.below(n - i));
            std::swap(all[i], all[j]);
        }
        cols.assign(all.begin(), all.begin() + k);
        std::sort(cols.begin(), cols.end());
        return;
    }
    // Sparse case: draw, sort, drop duplicates, redraw the shortfall.
    while (cols.size() < k) {
        const size_t need = k - cols.size();
        for (size_t i = 0; i < need; ++i) cols.push_back(uint32_t(r.below(n)));
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    }
}

}  // namespace

extern "C" int axspmv_gen_laplace2d(uint32_t grid, double diag, double off, axspmv_csr **out) {
    if (!out || grid == 0) return fail(AXSPMV_E_INVALID, "gen_laplace2d: grid must be > 0");
    if (uint64_t(grid) * grid * 5 >= (1ull << 32)) return fail(AXSPMV_E_OVERFLOW, "gen_laplace2d: grid too large");
    const uint32_t n = grid * grid;
    std::vector<uint32_t> rl(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = i / grid, b = i % grid;
        rl[i] = 1 + (a > 0) + (b > 0) + (b + 1 < grid) + (a + 1 < grid);
    }
    AX_TRY(alloc_from_lengths(n, n, rl, out));
    axspmv_csr *m = *out;
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = i / grid, b = i % grid;
        uint32_t p = m->row[i];
        auto put = [&](uint32_t c, double v) { m->col[p] = c; m->val[p] = v; ++p; };
        if (a > 0) put(i - grid, off);
        if (b > 0) put(i - 1, off);
        put(i, diag);
        if (b + 1 < grid) put(i + 1, off);
        if (a + 1 < grid) put(i + grid, off);
    }
    return axspmv_csr_stats(m);
}

extern "C" int axspmv_gen_powerlaw(uint32_t n, double alpha, uint32_t rlmax, double mean, uint64_t seed, axspmv_csr **out) {
    if (!out || n == 0 || rlmax == 0 || !(mean > 0)) return fail(AXSPMV_E_INVALID, "gen_powerlaw: bad arguments");
    rlmax = std::min(rlmax, n);
    // Exact inverse CDF of Zipf(alpha) on {1..rlmax}.
    std::vector<double> cdf(rlmax);
    double acc = 0.0;
    for (uint32_t k = 1; k <= rlmax; ++k) cdf[k - 1] = (acc += std::pow(double(k), -alpha));
    for (auto &c : cdf) c /= acc;
    cdf.back() = 1.0;
    std::vector<uint32_t> z(n);
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
        Rng r(seed, uint64_t(i) + 1);
        const double u = r.uniform();
        z[i] = uint32_t(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
        if (z[i] > rlmax) z[i] = rlmax;
    }
    // Scale s so that the clamped, rounded lengths average `mean`.
    auto lengths = [&](double s, std::vector<uint32_t> &rl) {
        double sum = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            double v = std::nearbyint(s * double(z[i]));
            v = std::min(std::max(v, 1.0), double(rlmax));
            rl[i] = uint32_t(v);
            sum += v;
        }
        return sum / double(n);
    };
    std::vector<uint32_t> rl(n);
    double zmean = 0.0;
    for (uint32_t v : z) zmean += v;
    zmean /= double(n);
    double s = mean / zmean;
    for (int it = 0; it < 6; ++it) s *= mean / lengths(s, rl);
    lengths(s, rl);
    AX_TRY(alloc_from_lengths(n, n, rl, out));
    axspmv_csr *m = *out;
    int bad = 0;
#pragma omp parallel
    {
        std::vector<uint32_t> cols;
#pragma omp for schedule(dynamic, 1024)
        for (uint32_t i = 0; i < n; ++i) {
            Rng r(seed, uint64_t(i) + 1);
            (void)r.uniform();  // the Zipf draw of pass 1
            distinct_sorted(r, n, rl[i], cols);
            const uint32_t p = m->row[i];
            for (uint32_t k = 0; k < rl[i]; ++k) {
                m->col[p + k] = cols[k];
                m->val[p + k] = r.normal();
            }
            if (cols.size() != rl[i]) bad = 1;
        }
    }
    if (bad) {
        axspmv_csr_free(m);
        *out = nullptr;
        return fail(AXSPMV_E_INVALID, "gen_powerlaw: column sampling failed");
    }
    return axspmv_csr_stats(m);
}

extern "C" int axspmv_gen_stencil27(uint32_t grid, double diag, double off, int permute, uint64_t seed, axspmv_csr **out) {
    if (!out || grid == 0) return fail(AXSPMV_E_INVALID, "gen_stencil27: grid must be > 0");
    const uint64_t n64 = uint64_t(grid) * grid * grid;
    if (n64 * 27 >= (1ull << 32)) return fail(AXSPMV_E_OVERFLOW, "gen_stencil27: grid too large");
    const uint32_t n = uint32_t(n64);
    auto cnt = [grid](uint32_t v) { return 1u + (v > 0) + (v + 1 < grid); };
    // Symmetric permutation P: new index of old row i is p[i] (Fisher-Yates).
    std::vector<uint32_t> p, q;
    if (permute) {
        p.resize(n);
        std::iota(p.begin(), p.end(), 0u);
        Rng r(seed, 0);
        for (uint32_t i = n - 1; i > 0; --i) std::swap(p[i], p[uint32_t(r.below(uint64_t(i) + 1))]);
        q.resize(n);
        for (uint32_t i = 0; i < n; ++i) q[p[i]] = i;
    }
    std::vector<uint32_t> rl(n);
#pragma omp parallel for schedule(static)
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t i = permute ? q[r] : r;
        const uint32_t a = i / (grid * grid), b = (i / grid) % grid, c = i % grid;
        rl[r] = cnt(a) * cnt(b) * cnt(c);
    }
    AX_TRY(alloc_from_lengths(n, n, rl, out));
    axspmv_csr *m = *out;
#pragma omp parallel
    {
        std::vector<std::pair<uint32_t, double>> ent;
#pragma omp for schedule(static)
        for (uint32_t r = 0; r < n; ++r) {
            const uint32_t i = permute ? q[r] : r;
            const int64_t a = i / (grid * grid), b = (i / grid) % grid, c = i % grid;
            ent.clear();
            for (int da = -1; da <= 1; ++da)
                for (int db = -1; db <= 1; ++db)
                    for (int dc = -1; dc <= 1; ++dc) {
                        const int64_t A = a + da, B = b + db, C = c + dc;
                        if (A < 0 || B < 0 || C < 0 || A >= grid || B >= grid || C >= grid) continue;
                        const uint32_t j = uint32_t((A * grid + B) * grid + C);
                        const uint32_t jc = permute ? p[j] : j;
                        ent.emplace_back(jc, (j == i) ? diag : off);
                    }
            std::sort(ent.begin(), ent.end());
            uint32_t pos = m->row[r];
            for (auto &e : ent) {
                m->col[pos] = e.first;
                m->val[pos] = e.second;
                ++pos;
            }
        }
    }
    return axspmv_csr_stats(m);
}

extern "C" int axspmv_gen_banded(uint32_t n, uint32_t hbw, double fill, uint64_t seed, axspmv_csr **out) {
    if (!out || n == 0 || !(fill >= 0 && fill <= 1)) return fail(AXSPMV_E_INVALID, "gen_banded: bad arguments");
    auto lo = [hbw](uint32_t i) { return i > hbw ? i - hbw : 0u; };
    auto hi = [hbw, n](uint32_t i) { return uint32_t(std::min<uint64_t>(uint64_t(i) + hbw, n - 1)); };
    std::vector<uint32_t> rl(n);
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
        Rng r(seed, uint64_t(i) + 1);
        uint32_t k = 0;
        for (uint32_t j = lo(i); j <= hi(i); ++j) k += r.uniform() < fill;
        rl[i] = k;
    }
    AX_TRY(alloc_from_lengths(n, n, rl, out));
    axspmv_csr *m = *out;
#pragma omp parallel
    {
        std::vector<uint32_t> cols;
#pragma omp for schedule(static)
        for (uint32_t i = 0; i < n; ++i) {
            Rng r(seed, uint64_t(i) + 1);
            cols.clear();
            for (uint32_t j = lo(i); j <= hi(i); ++j)
                if (r.uniform() < fill) cols.push_back(j);
            uint32_t p = m->row[i];
            for (uint32_t c : cols) {
                m->col[p] = c;
                m->val[p] = r.normal();
                ++p;
            }
        }
    }
    return axspmv_csr_stats(m);
}

extern "C" int axspmv_gen_random(uint32_t n, uint32_t rlmin, uint32_t rlmax, uint64_t seed, axspmv_csr **out) {
    if (!out || n == 0 || rlmin > rlmax) return fail(AXSPMV_E_INVALID, "gen_random: bad arguments");
    rlmax = std::min(rlmax, n);
    rlmin = std::min(rlmin, rlmax);
    std::vector<uint32_t> rl(n);
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
        Rng r(seed, uint64_t(i) + 1);
        rl[i] = rlmin + uint32_t(r.below(uint64_t(rlmax - rlmin) + 1));
    }
    AX_TRY(alloc_from_lengths(n, n, rl, out));
    axspmv_csr *m = *out;
#pragma omp parallel
    {
        std::vector<uint32_t> cols;
#pragma omp for schedule(static)
        for (uint32_t i = 0; i < n; ++i) {
            Rng r(seed, uint64_t(i) + 1);
            (void)r.below(uint64_t(rlmax - rlmin) + 1);
            distinct_sorted(r, n, rl[i], cols);
            const uint32_t p = m->row[i];
            for (uint32_t k = 0; k < rl[i]; ++k) {
                m->col[p + k] = cols[k];
                m->val[p + k] = r.normal();
            }
        }
    }
    return axspmv_csr_stats(m);
}

extern "C" int axspmv_fill_values_normal(axspmv_csr *a, double mu, double sigma, uint64_t seed) {
    if (!a || (a->nnz && !a->val)) return fail(AXSPMV_E_INVALID, "fill_values_normal: NULL matrix");
    Rng r(seed, 0);
    for (uint32_t k = 0; k < a->nnz; ++k) a->val[k] = mu + sigma * r.normal();
    return AXSPMV_OK;
}
