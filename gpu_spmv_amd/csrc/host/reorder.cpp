// Bandwidth-reducing symmetric reordering of a square matrix: reverse
// Cuthill-McKee (RCM) on the pattern of A + A^T, and the reordered matrix
// B = Q A Q^T.
//
// A random symmetric permutation (the 27-point stencil with permuted rows and
// columns) turns a banded matrix into one whose columns are spread over all
// of x; every gather then misses the L2s.  RCM numbers the vertices in
// breadth-first order from a vertex far from the rest of the graph, so that
// neighbours get nearby numbers: on a permuted grid the band comes back
// (width about twice the largest breadth-first level) and x is read through a
// narrow window again.  Plans built with AXSPMV_PLAN_REORDER hold B and run
// y = Q^T (B (Q x)).  No reference counterpart (the reference runs every
// matrix in the order it was read).
#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace axspmv {
namespace {

// Undirected adjacency of A + A^T without the diagonal: A's rows, and when A
// is not structurally symmetric also the rows of A^T's pattern.
struct Graph {
    uint32_t n = 0;
    const uint32_t *row = nullptr, *col = nullptr;  // A
    std::vector<uint32_t> trow, tcol;               // A^T's pattern, when needed
    std::vector<uint32_t> deg;
    template <class F>
    void each(uint32_t v, F &&f) const {
        for (uint32_t j = row[v]; j < row[v + 1]; ++j)
            if (col[j] != v) f(col[j]);
        if (!trow.empty())
            for (uint32_t j = trow[v]; j < trow[v + 1]; ++j)
                if (tcol[j] != v) f(tcol[j]);
    }
};

void build_graph(const axspmv_csr *a, Graph &g) {
    const uint32_t n = a->nrows;
    const uint32_t *row = a->row, *col = a->col;
    g.n = n;
    g.row = row;
    g.col = col;
    // Structurally symmetric A (every row's columns strictly increasing, and
    // every entry (i, j) with a partner (j, i), found by binary search in
    // row j): its rows are the graph.  Otherwise the rows of A^T's pattern
    // are added.
    bool sym = true;
#pragma omp parallel for schedule(dynamic, 4096) reduction(&& : sym)
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = row[i] + 1; j < row[i + 1]; ++j)
            if (col[j] <= col[j - 1]) {
                sym = false;
                break;
            }
    if (sym) {
#pragma omp parallel for schedule(dynamic, 4096) reduction(&& : sym)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j = row[i]; j < row[i + 1]; ++j)
                if (!std::binary_search(col + row[col[j]], col + row[col[j] + 1], i)) {
                    sym = false;
                    break;
                }
    }
    if (!sym) {
        // A^T's pattern, rows in no particular order (the numbering sorts
        // every vertex's new neighbours itself)
        std::vector<uint32_t> cnt(uint64_t(n) + 1, 0), tcol(a->nnz ? a->nnz : 1);
#pragma omp parallel for schedule(static)
        for (uint32_t j = 0; j < a->nnz; ++j) {
#pragma omp atomic
            cnt[col[j] + 1]++;
        }
        for (uint32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
        std::vector<uint32_t> next(cnt.begin(), cnt.end() - 1);
#pragma omp parallel for schedule(dynamic, 4096)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j = row[i]; j < row[i + 1]; ++j) {
                uint32_t at;
#pragma omp atomic capture
                at = next[col[j]]++;
                tcol[at] = i;
            }
        g.trow = std::move(cnt);
        g.tcol = std::move(tcol);
    }
    g.deg.assign(n, 0);
#pragma omp parallel for schedule(static)
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t d = 0;
        g.each(v, [&](uint32_t) { ++d; });
        g.deg[v] = d;
    }
}

// Breadth-first levels from r over the vertices not yet numbered (done[v] ==
// 0), stamping seen[v] = stamp; level by level, the vertices of a level
// shared out among the threads.  Returns the number of levels; `last` gets
// the vertices of the last level.
uint32_t bfs_levels(const Graph &g, uint32_t r, const std::vector<uint8_t> &done, std::vector<uint32_t> &seen,
                    uint32_t stamp, std::vector<uint32_t> &last) {
    std::vector<uint32_t> cur{r}, next;
    seen[r] = stamp;
    uint32_t levels = 1;
    for (;;) {
        next.clear();
        if (cur.size() < 4096) {  // small levels (small parts of the graph): one thread
            for (uint32_t v : cur)
                g.each(v, [&](uint32_t u) {
                    if (!done[u] && seen[u] != stamp) {
                        seen[u] = stamp;
                        next.push_back(u);
                    }
                });
        } else {
#pragma omp parallel
            {
                std::vector<uint32_t> mine;
#pragma omp for schedule(dynamic, 64) nowait
                for (size_t k = 0; k < cur.size(); ++k)
                    g.each(cur[k], [&](uint32_t u) {
                        if (done[u]) return;
                        uint32_t old = __atomic_load_n(&seen[u], __ATOMIC_RELAXED);
                        if (old != stamp && __atomic_compare_exchange_n(&seen[u], &old, stamp, false, __ATOMIC_RELAXED,
                                                                        __ATOMIC_RELAXED))
                            mine.push_back(u);
                    });
#pragma omp critical
                next.insert(next.end(), mine.begin(), mine.end());
            }
        }
        if (next.empty()) break;
        cur.swap(next);
        ++levels;
    }
    last.swap(cur);
    return levels;
}

}  // namespace

// q[k] = vertex numbered k.  Components are numbered one after another, each
// from a pseudo-peripheral vertex (George and Liu's search: breadth-first
// from a vertex of least degree, move to a least-degree vertex of the last
// level while that adds levels), neighbours in order of increasing degree
// (ties by index); the whole order is then reversed.
int rcm_order(const axspmv_csr *a, std::vector<uint32_t> &q) {
    AX_TRY(check_csr_bounds(a, "rcm_order"));
    if (a->nrows != a->ncols) return fail(AXSPMV_E_INVALID, "rcm_order: %u x %u matrix is not square", a->nrows, a->ncols);
    const uint32_t n = a->nrows;
    Graph g;
    build_graph(a, g);
    // vertices by increasing degree: the first one not yet numbered has the
    // least degree of its component
    std::vector<uint32_t> by_deg(n);
    std::iota(by_deg.begin(), by_deg.end(), 0u);
    std::stable_sort(by_deg.begin(), by_deg.end(), [&](uint32_t x, uint32_t y) { return g.deg[x] < g.deg[y]; });
    std::vector<uint8_t> done(n, 0);
    std::vector<uint32_t> seen(n, 0), last, nb;
    q.clear();
    q.reserve(n);
    uint32_t stamp = 0;
    for (uint32_t start : by_deg) {
        if (done[start]) continue;
        uint32_t r = start;
        uint32_t ecc = bfs_levels(g, r, done, seen, ++stamp, last);
        for (int it = 0; it < 16; ++it) {
            const uint32_t x = *std::min_element(last.begin(), last.end(), [&](uint32_t u, uint32_t v) {
                return g.deg[u] < g.deg[v] || (g.deg[u] == g.deg[v] && u < v);
            });
            std::vector<uint32_t> lx;
            const uint32_t ex = bfs_levels(g, x, done, seen, ++stamp, lx);
            if (ex <= ecc) break;
            r = x;
            ecc = ex;
            last.swap(lx);
        }
        // Cuthill-McKee from r
        const size_t base = q.size();
        q.push_back(r);
        done[r] = 1;
        for (size_t head = base; head < q.size(); ++head) {
            nb.clear();
            g.each(q[head], [&](uint32_t u) {
                if (!done[u]) {
                    done[u] = 1;
                    nb.push_back(u);
                }
            });
            std::sort(nb.begin(), nb.end(), [&](uint32_t u, uint32_t v) {
                return g.deg[u] < g.deg[v] || (g.deg[u] == g.deg[v] && u < v);
            });
            q.insert(q.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(q.begin(), q.end());
    return AXSPMV_OK;
}

// B = Q A Q^T: row k of B is row q[k] of A with every column c renumbered to
// qi[c] (qi the inverse of q) and the row's entries sorted by the new
// column; pos[e] = position in A of entry e of B.
int csr_permute_sym(const axspmv_csr *a, const std::vector<uint32_t> &q, axspmv_csr **out, std::vector<uint32_t> *pos) {
    AX_TRY(check_csr_bounds(a, "csr_permute"));
    if (!out) return fail(AXSPMV_E_INVALID, "csr_permute: NULL output");
    *out = nullptr;
    const uint32_t n = a->nrows;
    if (a->nrows != a->ncols) return fail(AXSPMV_E_INVALID, "csr_permute: %u x %u matrix is not square", a->nrows, a->ncols);
    if (q.size() != n) return fail(AXSPMV_E_INVALID, "csr_permute: %zu indices for %u rows", q.size(), n);
    std::vector<uint32_t> qi(n, 0xffffffffu);
    for (uint32_t k = 0; k < n; ++k) {
        if (q[k] >= n || qi[q[k]] != 0xffffffffu)
            return fail(AXSPMV_E_INVALID, "csr_permute: index %u at %u is out of range or repeated", q[k], k);
        qi[q[k]] = k;
    }
    axspmv_csr *b = nullptr;
    AX_TRY(axspmv_csr_alloc(n, n, a->nnz, &b));
    b->row[0] = 0;
    for (uint32_t k = 0; k < n; ++k) b->row[k + 1] = b->row[k] + (a->row[q[k] + 1] - a->row[q[k]]);
    if (pos) pos->assign(a->nnz, 0);
#pragma omp parallel
    {
        std::vector<std::pair<uint32_t, uint32_t>> e;  // (new column, position in A)
#pragma omp for schedule(dynamic, 1024)
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t i = q[k];
            e.clear();
            for (uint32_t j = a->row[i]; j < a->row[i + 1]; ++j) e.emplace_back(qi[a->col[j]], j);
            std::stable_sort(e.begin(), e.end(),
                             [](const std::pair<uint32_t, uint32_t> &x, const std::pair<uint32_t, uint32_t> &y) {
                                 return x.first < y.first;
                             });
            uint32_t o = b->row[k];
            for (const auto &p : e) {
                b->col[o] = p.first;
                b->val[o] = a->val[p.second];
                if (pos) (*pos)[o] = p.second;
                ++o;
            }
        }
    }
    *out = b;
    return AXSPMV_OK;
}

}  // namespace axspmv

extern "C" int axspmv_csr_rcm(const axspmv_csr *a, uint32_t *perm) {
    if (!a || (!perm && a->nrows)) return axspmv::fail(AXSPMV_E_INVALID, "csr_rcm: NULL argument");
    std::vector<uint32_t> q;
    AX_TRY(axspmv::rcm_order(a, q));
    std::copy(q.begin(), q.end(), perm);
    return AXSPMV_OK;
}

extern "C" int axspmv_csr_permute(const axspmv_csr *a, const uint32_t *perm, axspmv_csr **out) {
    if (!a || !out || (!perm && a->nrows)) return axspmv::fail(AXSPMV_E_INVALID, "csr_permute: NULL argument");
    std::vector<uint32_t> q(perm, perm + a->nrows);
    return axspmv::csr_permute_sym(a, q, out, nullptr);
}
