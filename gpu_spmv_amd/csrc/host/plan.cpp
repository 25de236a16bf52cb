// Device plans: upload a matrix in one format and run its kernel.
//
// Replaces the allocate / copy / launch / copy back / free sequence that
// every test_g*() of the reference repeats per call (cudaSpmv.cu:547-601,
// 871-929, 1100-1149, 1603-1663, 1963-2037, 2101-2170, 2213-2285): a plan
// uploads once, and axspmv_spmv() only launches.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

#include "../hip/launch.hpp"
#include "common.hpp"
#include "plan.hpp"

using namespace axspmv;

namespace axspmv {

int hip_fail(hipError_t e, const char *what) {
    return fail(AXSPMV_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace axspmv

void *axspmv_plan::dalloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        fail(AXSPMV_E_NOMEM, "hipMalloc of %zu bytes failed", bytes);
        return nullptr;
    }
    bufs.push_back(p);
    device_bytes += bytes;
    return p;
}

int axspmv_plan::upload(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return AXSPMV_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "plan upload");
}

axspmv_plan::~axspmv_plan() {
    delete T;
    delete Tdet;
    delete inner;
    for (void *p : bufs) (void)hipFree(p);
}

namespace axspmv {

// Greedy row-aligned wave blocks over units in row order: rows are packed
// into a block while the block holds at most `budget` units; a row longer
// than the budget gets a block of its own.  Long blocks go first, so the
// waves that loop longest start earliest.
std::vector<launch::Block> row_blocks(const std::vector<uint32_t> &units_per_row, uint32_t budget) {
    std::vector<launch::Block> small, big;
    uint64_t pos = 0;
    uint32_t beg = 0, cur = 0;
    bool open = false;
    for (uint32_t u : units_per_row) {
        if (u == 0) continue;
        if (u > budget) {
            if (open) small.push_back({beg, uint32_t(pos)});
            open = false;
            big.push_back({uint32_t(pos), uint32_t(pos + u)});
            pos += u;
            continue;
        }
        if (open && cur + u > budget) {
            small.push_back({beg, uint32_t(pos)});
            open = false;
        }
        if (!open) {
            beg = uint32_t(pos);
            cur = 0;
            open = true;
        }
        cur += u;
        pos += u;
    }
    if (open) small.push_back({beg, uint32_t(pos)});
    std::stable_sort(big.begin(), big.end(),
                     [](const launch::Block &a, const launch::Block &b) { return a.end - a.beg > b.end - b.beg; });
    big.insert(big.end(), small.begin(), small.end());
    return big;
}

std::vector<launch::Window> block_windows(const axspmv_csr *a, const std::vector<uint32_t> &units,
                                          const std::vector<launch::Block> &blocks) {
    const uint32_t n = a->nrows;
    std::vector<uint64_t> cum(size_t(n) + 1, 0);
    for (uint32_t r = 0; r < n; ++r) cum[r + 1] = cum[r] + units[r];
    std::vector<launch::Window> w(blocks.size(), launch::Window{0, 0});
    for (size_t i = 0; i < blocks.size(); ++i) {
        const launch::Block &b = blocks[i];
        uint32_t r = uint32_t(std::upper_bound(cum.begin(), cum.end(), uint64_t(b.beg)) - cum.begin()) - 1;
        uint32_t cmin = UINT32_MAX, cmax = 0;
        for (; r < n && cum[r] < b.end; ++r) {
            if (a->row[r + 1] == a->row[r]) continue;
            const auto mm = std::minmax_element(a->col + a->row[r], a->col + a->row[r + 1]);  // rows may be unsorted
            cmin = std::min(cmin, *mm.first);
            cmax = std::max(cmax, *mm.second);
        }
        if (cmin <= cmax) w[i] = launch::Window{cmin, cmax - cmin + 1};
    }
    return w;
}

bool worth_staging(const std::vector<launch::Window> &w) {
    size_t fit = 0;
    for (const launch::Window &x : w) fit += x.width <= launch::STAGE_COLS;
    return !w.empty() && 2 * fit >= w.size();
}

}  // namespace axspmv

namespace {

// Rows longer than this many slots are done one wave per row by SpMM.
constexpr uint32_t SPMM_LONG = 256;

// Row-contiguous layouts: record where each row starts (device array
// `first`, uploaded here unless it is the CSR row pointer itself) for SpMM
// and the value gradient.
int setup_rowwise(Plan &p, const axspmv_csr *a, const uint32_t *d_first, const void *d_val, const uint32_t *d_col,
                  hipStream_t s) {
    if (!p.d_rowptr) {
        if (p.d_row) {
            p.d_rowptr = p.d_row;
        } else {
            p.d_rowptr = static_cast<uint32_t *>(p.dalloc(4ull * (a->nrows + 1)));
            if (!p.d_rowptr) return AXSPMV_E_NOMEM;
            AX_TRY(p.upload(p.d_rowptr, a->row, 4ull * (a->nrows + 1), s));
        }
    }
    p.d_first = const_cast<uint32_t *>(d_first);
    std::vector<uint32_t> longrows;
    for (uint32_t r = 0; r < a->nrows; ++r)
        if (a->row[r + 1] - a->row[r] > SPMM_LONG) longrows.push_back(r);
    auto *lr = static_cast<uint32_t *>(p.dalloc(4 * longrows.size()));
    if (!lr) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(lr, longrows.data(), 4 * longrows.size(), s));
    p.spmm = launch::SpmmArgs{a->nrows, p.o.vtype, SPMM_LONG, p.d_first, p.d_rowptr, d_val, d_col, lr,
                              uint32_t(longrows.size())};
    p.rowwise = true;
    return AXSPMV_OK;
}

// Lanes per row for a row-per-group kernel: the power of two g in [2, 64]
// with 1.5 g >= the mean row length (CSR vector, AXC bricks).  Measured
// best or within 2 % of it on the 256^2 Laplacian, the 2^20-row power-law
// matrix and the permuted 27-point stencil for both; uniform banded rows of
// mean 19.5 prefer 8 lanes to the 16 it gives.
uint32_t lanes_for_mean(const axspmv_csr *a) {
    const double mean = a->nrows ? double(a->nnz) / a->nrows : 1.0;
    uint32_t g = 2;
    while (g < 64 && double(g) * 1.5 < mean) g *= 2;
    return g;
}

// True when the rows gather x from wide column ranges (not narrow bands).
bool wide_gathers(const axspmv_csr *a);

int plan_csr(Plan &p, const axspmv_csr *a, hipStream_t s) {
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    p.d_val = p.dalloc(vs * a->nnz);
    p.d_col = static_cast<uint32_t *>(p.dalloc(4ull * a->nnz));
    p.d_row = static_cast<uint32_t *>(p.dalloc(4ull * (a->nrows + 1)));
    if (!p.d_val || !p.d_col || !p.d_row) return AXSPMV_E_NOMEM;
    if (p.o.vtype == AXSPMV_F32) {
        std::vector<float> v(a->val, a->val + a->nnz);
        AX_TRY(p.upload(p.d_val, v.data(), 4ull * a->nnz, s));
    } else {
        AX_TRY(p.upload(p.d_val, a->val, 8ull * a->nnz, s));
    }
    AX_TRY(p.upload(p.d_col, a->col, 4ull * a->nnz, s));
    AX_TRY(p.upload(p.d_row, a->row, 4ull * (a->nrows + 1), s));
    p.slots = a->nnz;
    p.stream_bytes = (vs + 4) * a->nnz + 4ull * (a->nrows + 1);
    if (p.o.kernel == AXSPMV_KER_STREAM) {
        // lanes per row: opts.hbs, or from the mean row length
        const uint32_t g = p.o.hbs ? p.o.hbs : lanes_for_mean(a);
        if (!is_pow2(g) || g > 64) return fail(AXSPMV_E_UNSUPPORTED, "CSR vector: lanes per row %u not a power of two <= 64", g);
        p.lanes = g;
        // Non-temporal val / col loads for streams of more than twice the
        // Infinity Cache gathering wide: one box, the permuted 128^3 stencil
        // (669 MB) 797 -> 780 us; the 2^20-row power-law matrix (256 MB)
        // lost with them (259 -> 266 us), the stencil renumbered by reverse
        // Cuthill-McKee (narrow bands) stays level.
        p.seg_nt = p.stream_bytes > (512ull << 20) && wide_gathers(a) ? 1 : 0;
        p.kind = Plan::CSR_VECTOR;
        std::snprintf(p.name, sizeof p.name, "csr_vector_g%u", g);
        p.launches = 2;
        return setup_rowwise(p, a, p.d_row, p.d_val, p.d_col, s);
    }
    p.kind = Plan::CSR_NAIVE;
    std::snprintf(p.name, sizeof p.name, "csr_naive");
    return setup_rowwise(p, a, p.d_row, p.d_val, p.d_col, s);
}

int upload_planes(Plan &p, const double *vplane, const uint32_t *cplane, uint64_t nslots, hipStream_t s) {
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    p.d_vplane = p.dalloc(vs * nslots);
    p.d_cplane = static_cast<uint32_t *>(p.dalloc(4 * nslots));
    if (!p.d_vplane || !p.d_cplane) return AXSPMV_E_NOMEM;
    if (p.o.vtype == AXSPMV_F32) {
        std::vector<float> v(vplane, vplane + nslots);
        AX_TRY(p.upload(p.d_vplane, v.data(), 4 * nslots, s));
    } else {
        AX_TRY(p.upload(p.d_vplane, vplane, 8 * nslots, s));
    }
    AX_TRY(p.upload(p.d_cplane, cplane, 4 * nslots, s));
    p.slots = nslots;
    return AXSPMV_OK;
}

int upload_blocks(Plan &p, const std::vector<launch::Block> &blocks, const std::vector<uint32_t> &empty, hipStream_t s) {
    p.nblocks = uint32_t(blocks.size());
    p.d_blocks = static_cast<launch::Block *>(p.dalloc(sizeof(launch::Block) * blocks.size()));
    p.nempty = uint32_t(empty.size());
    p.d_empty = static_cast<uint32_t *>(p.dalloc(4 * empty.size()));
    if (!p.d_blocks || !p.d_empty) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(p.d_blocks, blocks.data(), sizeof(launch::Block) * blocks.size(), s));
    AX_TRY(p.upload(p.d_empty, empty.data(), 4 * empty.size(), s));
    return AXSPMV_OK;
}

// Row of CSR position q (q < nnz).
uint32_t row_of(const axspmv_csr *a, uint32_t q) {
    return uint32_t(std::upper_bound(a->row, a->row + a->nrows + 1, q) - a->row) - 1;
}

int plan_axt_com(Plan &p, const axspmv_csr *a, hipStream_t s) {
    const uint32_t thw = p.o.thw, th = p.o.th;
    axspmv_axt *m = nullptr;
    AX_TRY(axspmv_build_axt(a, nullptr, p.o.bs, thw, th, AXSPMV_AXT_COM, AXSPMV_BUILD_PLANES, &m, nullptr));
    std::unique_ptr<axspmv_axt, void (*)(axspmv_axt *)> guard(m, axspmv_axt_free);
    AX_TRY(upload_planes(p, m->vplane, m->cplane, uint64_t(m->lenAX) / 2, s));
    p.d_sec = static_cast<uint32_t *>(p.dalloc(4ull * m->lenSEC));
    if (!p.d_sec) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(p.d_sec, m->sec, 4ull * m->lenSEC, s));
    const uint32_t nnz = a->nnz;
    // Wave ranges: th == 1 in slots, a multiple of bs and of 256 (4 slots x
    // 64 lanes); th > 1 in tile columns, 4 passes of 64 columns.  A small
    // matrix gets shorter ranges (down to one pass) until about 2048 waves
    // spread over the 1024 SIMDs.
    const uint32_t units = th == 1 ? nnz : uint32_t(ceil_div(nnz, th));
    uint32_t wr = th == 1 ? std::max<uint32_t>(p.o.bs, 1024) : 256;
    const uint32_t wr_min = th == 1 ? std::max<uint32_t>(p.o.bs, 256) : 64;
    while (wr > wr_min && units / wr < 2048) wr /= 2;
    std::vector<launch::ComWave> waves;
    for (uint64_t b = 0; b < units; b += wr) {
        launch::ComWave w{uint32_t(b), uint32_t(std::min<uint64_t>(b + wr, units)), 0, 0};
        const uint64_t S = th == 1 ? w.beg : uint64_t(w.beg) * th;
        const uint64_t E = th == 1 ? w.end : std::min<uint64_t>(uint64_t(w.end) * th, nnz);
        if (S > 0 && row_of(a, uint32_t(S)) == row_of(a, uint32_t(S - 1))) w.flags |= 1;
        if (E < nnz && row_of(a, uint32_t(E - 1)) == row_of(a, uint32_t(E))) w.flags |= 2;
        waves.push_back(w);
    }
    // Rows spanning waves: (row, first wave, last wave).
    std::vector<uint32_t> span;
    bool open = false;
    uint32_t orow = 0, w0 = 0;
    for (uint32_t k = 0; k < waves.size(); ++k) {
        const uint64_t E = th == 1 ? waves[k].end : std::min<uint64_t>(uint64_t(waves[k].end) * th, nnz);
        const bool out = waves[k].flags & 2;
        const uint32_t lastrow = row_of(a, uint32_t(E - 1));
        if (open && !(out && lastrow == orow)) {
            span.insert(span.end(), {orow, w0, k});
            open = false;
        }
        if (out && !open) {
            open = true;
            orow = lastrow;
            w0 = k;
        }
    }
    std::vector<uint32_t> empty;
    for (uint32_t r = 0; r < a->nrows; ++r)
        if (a->row[r + 1] == a->row[r]) empty.push_back(r);
    auto *dw = static_cast<launch::ComWave *>(p.dalloc(sizeof(launch::ComWave) * waves.size()));
    auto *dspan = static_cast<uint32_t *>(p.dalloc(4 * span.size()));
    auto *dempty = static_cast<uint32_t *>(p.dalloc(4 * empty.size()));
    auto *head = static_cast<double *>(p.dalloc(8 * waves.size()));
    auto *tail = static_cast<double *>(p.dalloc(8 * waves.size()));
    if (!dw || !dspan || !dempty || !head || !tail) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(dw, waves.data(), sizeof(launch::ComWave) * waves.size(), s));
    AX_TRY(p.upload(dspan, span.data(), 4 * span.size(), s));
    AX_TRY(p.upload(dempty, empty.data(), 4 * empty.size(), s));
    launch::ComArgs c{};
    c.vplane = p.d_vplane;
    c.cplane = p.d_cplane;
    c.hdr = p.d_sec;
    c.waves = dw;
    c.nwaves = uint32_t(waves.size());
    c.thw = thw;
    c.th = th;
    c.bs = p.o.bs;
    c.log = m->log;
    c.nnz = nnz;
    c.vtype = p.o.vtype;
    c.head = head;
    c.tail = tail;
    c.span = dspan;
    c.nspan = uint32_t(span.size() / 3);
    c.empty = dempty;
    c.nempty = uint32_t(empty.size());
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    p.stream_bytes = (vs + 8) * p.slots + sizeof(launch::ComWave) * waves.size();
    // Non-temporal plane loads where the stream (more than 128 MB) would
    // share the Infinity Cache with the x lines the gathers reuse; th = 1 on
    // narrow bands keeps ordinary loads, as AXT UNC th = 1 does.  One box, per
    // SpMV, th = 1 / th = 4: the 2^20-row power-law matrix 271.2 -> 255.4 /
    // 251.5 -> 235.1 us, the permuted 128^3 stencil 867 -> 839.5 / 826 ->
    // 803 us, the same renumbered by reverse Cuthill-McKee 306.5 -> 311.5
    // (hence the exception) / 281.5 -> 266.3 us.
    c.nt = p.stream_bytes > (128ull << 20) && (th > 1 || wide_gathers(a)) ? 1 : 0;
    p.com = c;
    p.nblocks = c.nwaves;
    p.launches = 1 + (c.nspan > 0) + (c.nempty > 0);
    p.kind = Plan::AXT_COM;
    if (th == 1)
        std::snprintf(p.name, sizeof p.name, "axt_com_h1_thw%u_bs%u", thw, p.o.bs);
    else
        std::snprintf(p.name, sizeof p.name, "axt_com_h%u_thw%u", th, thw);
    return AXSPMV_OK;
}

// The reference's interleaved ax of AXT UNC th == 1 (values, x-halves left
// zero: every SpMV refreshes them) and the column of every value slot, with
// 0xffffffff on padding slots so the refresh writes 0 there as the builder.
int upload_ax_layout(Plan &p, const axspmv_csr *a, const axspmv_axt *m, hipStream_t s) {
    const uint32_t thw = m->tileHW;
    const uint64_t ns = uint64_t(m->lenAX) / 2;
    std::vector<double> ax(2 * ns, 0.0);
    std::vector<uint32_t> map(ns ? ns : 1, 0xffffffffu);
    for (uint64_t q = 0; q < ns; ++q) ax[(q / thw) * 2 * thw + q % thw] = m->vplane[q];
    for (uint32_t r = 0; r < a->nrows; ++r) {
        const uint64_t first = m->con[r] / 2, rl = a->row[r + 1] - a->row[r];
        for (uint64_t k = 0; k < rl; ++k) map[first + k] = m->cplane[first + k];
    }
    p.d_ax = static_cast<double *>(p.dalloc(16 * ns));
    p.d_axmap = static_cast<uint32_t *>(p.dalloc(4 * ns));
    if (!p.d_ax || !p.d_axmap) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(p.d_ax, ax.data(), 16 * ns, s));
    AX_TRY(p.upload(p.d_axmap, map.data(), 4 * ns, s));
    p.slots = ns;
    return AXSPMV_OK;
}

// AXT UNC th == 1 in S column slices: slice k holds the entries of columns
// [k C, (k + 1) C) (C = ceil(ncols / S)) and is the AXT layout of that part
// of the matrix, built by the same builder (global column indices, so x is
// indexed directly).  The slices' planes and sec arrays are concatenated;
// every slice gets its own row-aligned wave blocks (unit offsets moved to the
// concatenated planes).  The dispatch list gives workgroup g (four waves,
// four blocks) slice (g % 8) % S, and workgroups run on XCD g % 8, so an XCD
// only gathers from its slice of x -- 2 MB at S = 4 on the 2^20-row matrix,
// resident in its 4 MB L2.  Rows accumulate across slices with FP64 atomic
// adds into a y cleared first.
bool narrow_bands(const axspmv_csr *a, uint32_t rows, uint64_t cols);

// AXT UNC streams its planes with non-temporal loads where the stream is
// larger than the 256 MB Infinity Cache (it comes from HBM on every SpMV
// anyway) and the rows gather x from wide column ranges (so the x lines the
// gathers share are what the cache should keep).  One box, per SpMV: the
// 2^20-row power-law matrix th = 1 218.7 -> 209.2 us (4 column slices 168.5
// -> 157.8), th = 4 / 8 226.0 -> 218.3 / 238.0 -> 227.8 us; the permuted
// 128^3 stencil th = 1 806.7 -> 784.5 us (slices 452.5 -> 412.0), th = 4 /
// 8 805.4 -> 798.4 / 822.8 -> 813.0 us; the stencil renumbered by reverse
// Cuthill-McKee, whose rows read narrow bands, lost with them (th = 1:
// 197.3 -> 204.2 us).  A 128 MB bound (K1's) would take in the power-law
// matrix with FP32 values (188 MB): th = 1 196.7 -> 192.5 us, but its
// 4-slice plan 123.0 -> 146.0 us.
bool wide_gathers(const axspmv_csr *a) {
    const uint32_t rows = uint32_t(std::min<uint64_t>(16384, ceil_div(ceil_div(uint64_t(a->nrows), 256), 64) * 64));
    return !narrow_bands(a, std::max<uint32_t>(64, rows), (2ull << 20) / 8);
}
uint32_t axt_stream_nt(const axspmv_csr *a, uint64_t stream_bytes) {
    return stream_bytes > (256ull << 20) && wide_gathers(a) ? 1 : 0;
}

int plan_axt_sliced(Plan &p, const axspmv_csr *a, hipStream_t s) {
    const uint32_t thw = p.o.thw, S = p.o.slices, n = a->nrows;
    if (S != 2 && S != 4 && S != 8) return fail(AXSPMV_E_UNSUPPORTED, "AXT column slices must be 2, 4 or 8, got %u", S);
    const uint64_t C = ceil_div(std::max<uint32_t>(a->ncols, 1), S);
    const uint32_t K = std::min<uint32_t>(4, thw);
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    // entries of each row by slice
    std::vector<uint32_t> cnt(size_t(S) * n, 0);
    for (uint32_t r = 0; r < n; ++r)
        for (uint32_t j = a->row[r]; j < a->row[r + 1]; ++j) cnt[size_t(a->col[j] / C) * n + r]++;
    std::vector<double> vcat;
    std::vector<uint32_t> ccat, scat;
    std::vector<std::vector<launch::Block>> sblocks(S);
    uint64_t units_before = 0;
    for (uint32_t k = 0; k < S; ++k) {
        axspmv_csr sub{};
        sub.nrows = n;
        sub.ncols = a->ncols;
        std::vector<uint32_t> row(size_t(n) + 1, 0), col;
        std::vector<double> val;
        for (uint32_t r = 0; r < n; ++r) row[r + 1] = row[r] + cnt[size_t(k) * n + r];
        col.resize(row[n]);
        val.resize(row[n]);
        for (uint32_t r = 0; r < n; ++r) {
            uint32_t q = row[r];
            for (uint32_t j = a->row[r]; j < a->row[r + 1]; ++j)
                if (a->col[j] / C == k) {
                    col[q] = a->col[j];
                    val[q++] = a->val[j];
                }
        }
        sub.nnz = row[n];
        sub.row = row.data();
        sub.col = col.data();
        sub.val = val.data();
        axspmv_axt *m = nullptr;
        AX_TRY(axspmv_build_axt(&sub, nullptr, p.o.bs, thw, 1, AXSPMV_AXT_UNC,
                               AXSPMV_BUILD_PLANES | AXSPMV_BUILD_PAD_COL, &m, nullptr));
        std::unique_ptr<axspmv_axt, void (*)(axspmv_axt *)> guard(m, axspmv_axt_free);
        if (vcat.size() + m->lenAX / 2 >= (1ull << 32))
            return fail(AXSPMV_E_OVERFLOW, "AXT slices: %llu slots exceed 32 bits",
                        (unsigned long long)(vcat.size() + m->lenAX / 2));
        vcat.insert(vcat.end(), m->vplane, m->vplane + m->lenAX / 2);
        ccat.insert(ccat.end(), m->cplane, m->cplane + m->lenAX / 2);
        scat.insert(scat.end(), m->sec, m->sec + m->lenSEC);
        std::vector<uint32_t> units(n);
        for (uint32_t r = 0; r < n; ++r) units[r] = uint32_t(ceil_div(cnt[size_t(k) * n + r], thw) * (thw / K));
        sblocks[k] = row_blocks(units, p.o.budget ? p.o.budget : 64);
        for (launch::Block &b : sblocks[k]) {
            b.beg += uint32_t(units_before);
            b.end += uint32_t(units_before);
        }
        units_before += uint64_t(m->lenAX / 2) / K;
    }
    // dispatch list: workgroup g takes blocks 4i .. 4i + 3 of slice (g % 8) % S,
    // i = (g / 8) * (8 / S) + (g % 8) / S; missing blocks are empty
    size_t most = 0;
    for (const auto &b : sblocks) most = std::max(most, b.size());
    const uint32_t per = 8 / S;
    const uint64_t groups = ceil_div(ceil_div(most, 4), per) * 8;
    std::vector<launch::Block> list(groups * 4, launch::Block{0, 0});
    for (uint64_t g = 0; g < groups; ++g) {
        const uint32_t xcd = uint32_t(g % 8), k = xcd % S;
        const uint64_t i = (g / 8) * per + xcd / S;
        for (uint32_t w = 0; w < 4; ++w)
            if (4 * i + w < sblocks[k].size()) list[4 * g + w] = sblocks[k][4 * i + w];
    }
    if (list.size() >= (1ull << 32) / 64) return fail(AXSPMV_E_OVERFLOW, "AXT slices: too many wave blocks");
    AX_TRY(upload_planes(p, vcat.data(), ccat.data(), vcat.size(), s));
    p.d_sec = static_cast<uint32_t *>(p.dalloc(4ull * scat.size()));
    if (!p.d_sec) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(p.d_sec, scat.data(), 4ull * scat.size(), s));
    // Deterministic: slice partials stored in part[k * n + r] (0 for the rows
    // without entries in slice k), summed in slice order by a merge kernel.
    const bool det = p.o.flags & AXSPMV_PLAN_DETERMINISTIC;
    std::vector<uint32_t> empty;
    if (det) {
        if (uint64_t(S) * n >= (1ull << 32)) return fail(AXSPMV_E_OVERFLOW, "AXT slices: S x nrows exceeds 32 bits");
        for (uint32_t k = 0; k < S; ++k)
            for (uint32_t r = 0; r < n; ++r)
                if (!cnt[size_t(k) * n + r]) empty.push_back(k * n + r);
        p.d_part = static_cast<double *>(p.dalloc(8ull * S * n));
        if (!p.d_part) return AXSPMV_E_NOMEM;
    }
    AX_TRY(upload_blocks(p, list, empty, s));
    p.thw = thw;
    p.th = 1;
    p.slices = S;
    p.slots = vcat.size();
    p.stream_bytes = (vs + 4) * p.slots + 4ull * scat.size() + 8ull * list.size();
    p.seg_nt = axt_stream_nt(a, p.stream_bytes);
    p.launches = 2;
    p.kind = Plan::AXT_UNC_H1_SLICED;
    std::snprintf(p.name, sizeof p.name, "axt_unc_h1_thw%u_s%u%s", thw, S, det ? "_det" : "");
    return AXSPMV_OK;
}

int plan_axt(Plan &p, const axspmv_csr *a, hipStream_t s) {
    const uint32_t thw = p.o.thw, th = p.o.th;
    if (!is_pow2(thw) || thw > 64) return fail(AXSPMV_E_UNSUPPORTED, "AXT kernels need thw a power of two <= 64, got %u", thw);
    const bool axl = p.o.flags & AXSPMV_PLAN_AX_LAYOUT;
    if (axl && (th != 1 || p.o.mode != AXSPMV_AXT_UNC || p.o.vtype != AXSPMV_F64))
        return fail(AXSPMV_E_UNSUPPORTED, "AX layout: AXT UNC th = 1 with FP64 values only");
    if (p.o.slices > 1) {
        if (th != 1 || p.o.mode != AXSPMV_AXT_UNC || axl || (p.o.flags & AXSPMV_PLAN_STAGE_X))
            return fail(AXSPMV_E_UNSUPPORTED, "AXT column slices: UNC th = 1 only, without the AX layout or x staging");
        return plan_axt_sliced(p, a, s);
    }
    if (p.o.mode == AXSPMV_AXT_COM) return plan_axt_com(p, a, s);
    axspmv_axt *m = nullptr;
    AX_TRY(axspmv_build_axt(a, nullptr, p.o.bs, thw, th, p.o.mode, AXSPMV_BUILD_PLANES | AXSPMV_BUILD_PAD_COL, &m,
                           nullptr));
    std::unique_ptr<axspmv_axt, void (*)(axspmv_axt *)> guard(m, axspmv_axt_free);
    if (axl)
        AX_TRY(upload_ax_layout(p, a, m, s));
    else
        AX_TRY(upload_planes(p, m->vplane, m->cplane, uint64_t(m->lenAX) / 2, s));
    p.d_sec = static_cast<uint32_t *>(p.dalloc(4ull * m->lenSEC));
    if (!p.d_sec) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(p.d_sec, m->sec, 4ull * m->lenSEC, s));
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    std::vector<uint32_t> units(a->nrows), empty;
    const uint32_t K = std::min<uint32_t>(4, thw);
    for (uint32_t r = 0; r < a->nrows; ++r) {
        const uint32_t rl = a->row[r + 1] - a->row[r];
        if (rl == 0) empty.push_back(r);
        units[r] = th == 1 ? uint32_t(ceil_div(rl, thw) * (thw / K)) : uint32_t(ceil_div(rl, th));
    }
    // Budget in units per wave block: 64 (one pass of the wave) by default.
    // Measured against 256 (four passes): the 2^20-row power-law matrix th =
    // 1 / 4 / 8: 218 / 225 / 234 us against 227 / 240 / 271 us; the permuted
    // 27-point stencil on 128^3 th = 1 / 4: 801 / 810 against 812 / 839 us;
    // the 256^2 Laplacian (graph-timed) thw 4: 2.6 against 4.1 us.  Shorter
    // blocks balance better and, on small matrices, spread over the 1024
    // SIMDs; a banded matrix at th = 4 is the exception (5 % slower).
    const uint32_t budget = p.o.budget ? p.o.budget : 64;
    const std::vector<launch::Block> blocks = row_blocks(units, budget);
    AX_TRY(upload_blocks(p, blocks, empty, s));
    // On request, banded / stencil structure: blocks whose columns fit a
    // per-wave LDS window read x from there (not the AX layout, which reads
    // no x).
    if (!axl && (p.o.flags & AXSPMV_PLAN_STAGE_X)) {
        const std::vector<launch::Window> win = block_windows(a, units, blocks);
        if (worth_staging(win)) {
            for (const launch::Window &w : win) p.staged_blocks += w.width <= launch::STAGE_COLS;
            p.d_win = static_cast<launch::Window *>(p.dalloc(sizeof(launch::Window) * win.size()));
            if (!p.d_win) return AXSPMV_E_NOMEM;
            AX_TRY(p.upload(p.d_win, win.data(), sizeof(launch::Window) * win.size(), s));
        }
    }
    p.thw = thw;
    p.th = th;
    if (axl) {
        // refresh: map (4 B) in, x-half (8 B) out per slot; SpMV: 16 B per slot
        p.stream_bytes = 28ull * p.slots + 4ull * m->lenSEC + 8ull * p.nblocks;
        p.launches = 2;
        p.kind = Plan::AXT_UNC_H1_AX;
        std::snprintf(p.name, sizeof p.name, "axt_unc_h1_thw%u_ax", thw);
        return AXSPMV_OK;
    }
    p.stream_bytes = (vs + 4) * p.slots + 4ull * m->lenSEC + 8ull * p.nblocks;
    if (th != 1) {
        p.seg_nt = axt_stream_nt(a, p.stream_bytes);
        p.kind = Plan::AXT_UNC_H;
        std::snprintf(p.name, sizeof p.name, "axt_unc_h%u_thw%u", th, thw);
        return AXSPMV_OK;
    }
    p.kind = Plan::AXT_UNC_H1;
    p.seg_nt = axt_stream_nt(a, p.stream_bytes);
    std::snprintf(p.name, sizeof p.name, "axt_unc_h1_thw%u", thw);
    // Rows are contiguous in the planes: first slot of row r is con[r]/2.
    std::vector<uint32_t> first(a->nrows ? a->nrows : 1);
    for (uint32_t r = 0; r < a->nrows; ++r) first[r] = m->con[r] / 2;
    auto *d_first = static_cast<uint32_t *>(p.dalloc(4ull * first.size()));
    if (!d_first) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(d_first, first.data(), 4ull * first.size(), s));
    p.nunits = uint32_t(uint64_t(m->tileN) * thw / K);
    // One-pass blocks (<= 64 units) for the fused forward/adjoint kernel.
    const std::vector<launch::Block> fb = row_blocks(units, 64);
    p.nfblocks = uint32_t(fb.size());
    p.d_fblocks = static_cast<launch::Block *>(p.dalloc(sizeof(launch::Block) * fb.size()));
    if (!p.d_fblocks) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(p.d_fblocks, fb.data(), sizeof(launch::Block) * fb.size(), s));
    return setup_rowwise(p, a, d_first, p.d_vplane, p.d_cplane, s);
}

// Steps of one 64-lane K1 work item at most; longer chunks are split.
constexpr uint32_t K1_STEPS = 64;

int plan_k1(Plan &p, const axspmv_csr *a, hipStream_t s) {
    const uint32_t C = p.o.chunk;
    if (!is_pow2(C) || C > 64) return fail(AXSPMV_E_UNSUPPORTED, "K1 kernel needs chunk a power of two <= 64, got %u", C);
    axspmv_k1 *m = nullptr;
    AX_TRY(axspmv_build_k1(a, C, p.o.reorder ? 1 : 0, &m, nullptr));
    std::unique_ptr<axspmv_k1, void (*)(axspmv_k1 *)> guard(m, axspmv_k1_free);
    // Padding slots (steps past a row's length, and the lanes past the last
    // row) get the column AXSPMV_PAD_COL, which the kernel reads as 0.0.
#pragma omp parallel for schedule(static)
    for (uint32_t c = 0; c < m->chunkNum; ++c)
        for (uint32_t r = 0; r < C; ++r) {
            const uint64_t i = uint64_t(c) * C + r;
            const uint32_t len = i < a->nrows ? a->row[m->permi[i] + 1] - a->row[m->permi[i]] : 0;
            for (uint32_t k = len; k < m->nmc[c]; ++k) m->col[uint64_t(m->chp[c]) + uint64_t(k) * C + r] = AXSPMV_PAD_COL;
        }
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    std::vector<launch::K1Item> items;
    std::vector<uint32_t> split;
    uint32_t nslot = 0;
    for (uint32_t c = 0; c < m->chunkNum; ++c) {
        const uint32_t nm = m->nmc[c];
        if (nm <= K1_STEPS) {
            items.push_back({c, m->chp[c], 0, nm, 0xffffffffu});  // writes y directly
            continue;
        }
        const uint32_t parts = uint32_t(ceil_div(nm, K1_STEPS));
        split.insert(split.end(), {c, nslot, parts});
        for (uint32_t k = 0; k < parts; ++k)
            items.push_back({c, m->chp[c], k * K1_STEPS, std::min(nm, (k + 1) * K1_STEPS), nslot + k});
        nslot += parts;
    }
    void *val = p.dalloc(vs * m->lenVC);
    auto *col = static_cast<uint32_t *>(p.dalloc(4ull * m->lenVC));
    auto *permi = static_cast<uint32_t *>(p.dalloc(4ull * (m->chunkNum + 1) * C));
    auto *it = static_cast<launch::K1Item *>(p.dalloc(sizeof(launch::K1Item) * items.size()));
    auto *sp = static_cast<uint32_t *>(p.dalloc(4 * split.size()));
    auto *part = static_cast<double *>(p.dalloc(8ull * nslot * C));
    if (!val || !col || !permi || !it || !sp || !part) return AXSPMV_E_NOMEM;
    if (p.o.vtype == AXSPMV_F32) {
        std::vector<float> v(m->val, m->val + m->lenVC);
        AX_TRY(p.upload(val, v.data(), 4ull * m->lenVC, s));
    } else {
        AX_TRY(p.upload(val, m->val, 8ull * m->lenVC, s));
    }
    AX_TRY(p.upload(col, m->col, 4ull * m->lenVC, s));
    AX_TRY(p.upload(permi, m->permi, 4ull * (m->chunkNum + 1) * C, s));
    AX_TRY(p.upload(it, items.data(), sizeof(launch::K1Item) * items.size(), s));
    AX_TRY(p.upload(sp, split.data(), 4 * split.size(), s));
    p.k1 = launch::K1Args{it, uint32_t(items.size()), sp, uint32_t(split.size() / 3), C, a->nrows, p.o.vtype,
                          val, col, permi, part};
    p.slots = m->lenVC;
    p.stream_bytes = (vs + 4) * uint64_t(m->lenVC) + 4ull * a->nrows + sizeof(launch::K1Item) * items.size();
    // A stream of more than 128 MB shares the Infinity Cache with the x
    // lines the gathers reuse (a chunk's lanes are rows of one length, not
    // neighbours): non-temporal loads keep the cache for x.  One box, per
    // SpMV: the 2^20-row power-law matrix 226.5 -> 214.4 us, the permuted
    // 128^3 stencil 740 -> 719.6 us, the same renumbered by reverse
    // Cuthill-McKee 209.2 -> 193.2 us.
    p.k1.nt = p.stream_bytes > (128ull << 20) ? 1 : 0;
    p.launches = split.empty() ? 1 : 2;
    p.nblocks = uint32_t(items.size());
    p.kind = Plan::K1;
    std::snprintf(p.name, sizeof p.name, "k1_c%u%s", C, p.o.reorder ? "" : "_natural");
    return AXSPMV_OK;
}

int plan_axc(Plan &p, const axspmv_csr *a, hipStream_t s) {
    if (p.o.hbs == 0) p.o.hbs = lanes_for_mean(a);  // bricks (and lanes per row) from the mean row length
    const uint32_t G = p.o.hbs;
    if (!is_pow2(G) || G > 64) return fail(AXSPMV_E_UNSUPPORTED, "AXC kernel needs hbs a power of two <= 64, got %u", G);
    if (p.o.flags & AXSPMV_PLAN_AX_LAYOUT) {
        // AXC's ax (bricks of hbs values then hbs x values from brp[r],
        // get_axAXC cudaSpmv.cu:952-979) is the AXT UNC th = 1 layout with
        // thw = hbs, slot for slot: the same plan serves it.
        p.o.thw = G;
        p.o.th = 1;
        p.o.mode = AXSPMV_AXT_UNC;
        AX_TRY(plan_axt(p, a, s));
        std::snprintf(p.name, sizeof p.name, "axc_hbs%u_ax", G);
        return AXSPMV_OK;
    }
    axspmv_axc *m = nullptr;
    AX_TRY(axspmv_build_axc(a, nullptr, G, AXSPMV_BUILD_PLANES | AXSPMV_BUILD_PAD_COL, &m, nullptr));
    std::unique_ptr<axspmv_axc, void (*)(axspmv_axc *)> guard(m, axspmv_axc_free);
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    const uint32_t long_slots = std::max<uint32_t>(64 * G, 256);
    std::vector<uint32_t> longrows;
    for (uint32_t r = 0; r < a->nrows; ++r)
        if ((m->brp[r + 1] - m->brp[r]) / 2 > long_slots) longrows.push_back(r);
    const uint64_t nslots = m->lenAX / 2;
    void *vp = p.dalloc(vs * nslots);
    auto *mapx = static_cast<uint32_t *>(p.dalloc(4ull * m->lenMAPX));
    auto *brp = static_cast<uint32_t *>(p.dalloc(4ull * m->lenBRP));
    auto *lr = static_cast<uint32_t *>(p.dalloc(4 * longrows.size()));
    if (!vp || !mapx || !brp || !lr) return AXSPMV_E_NOMEM;
    if (p.o.vtype == AXSPMV_F32) {
        std::vector<float> v(m->vplane, m->vplane + nslots);
        AX_TRY(p.upload(vp, v.data(), 4 * nslots, s));
    } else {
        AX_TRY(p.upload(vp, m->vplane, 8 * nslots, s));
    }
    AX_TRY(p.upload(mapx, m->mapx, 4ull * m->lenMAPX, s));
    AX_TRY(p.upload(brp, m->brp, 4ull * m->lenBRP, s));
    AX_TRY(p.upload(lr, longrows.data(), 4 * longrows.size(), s));
    p.axc = launch::AxcArgs{a->nrows, G, p.o.vtype, long_slots, brp, mapx, vp, lr, uint32_t(longrows.size())};
    p.slots = nslots;
    p.stream_bytes = (vs + 4) * nslots + 4ull * m->lenBRP;
    // Non-temporal plane loads for streams of more than twice the Infinity
    // Cache gathering wide: one box, the permuted 128^3 stencil hbs 4 / 64
    // 1214.7 -> 1163.0 / 945.7 -> 914.3 us; on the 2^20-row power-law matrix
    // (streams of 300-400 MB) hbs 4 lost with them (301 -> 347 us) where hbs
    // 64 gained (400 -> 384 us); renumbered (narrow bands) level.
    p.axc.nt = p.stream_bytes > (512ull << 20) && wide_gathers(a) ? 1 : 0;
    p.launches = longrows.empty() ? 1 : 2;
    p.kind = Plan::AXC;
    std::snprintf(p.name, sizeof p.name, "axc_hbs%u", G);
    // Rows are contiguous in the value plane: first slot of row r is brp[r]/2.
    std::vector<uint32_t> first(a->nrows ? a->nrows : 1);
    for (uint32_t r = 0; r < a->nrows; ++r) first[r] = m->brp[r] / 2;
    auto *d_first = static_cast<uint32_t *>(p.dalloc(4ull * first.size()));
    if (!d_first) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(d_first, first.data(), 4ull * first.size(), s));
    return setup_rowwise(p, a, d_first, vp, mapx, s);
}

// Row-block starts for RCT: nrb blocks of about nnz / nrb nonzeros each,
// every block 1..rmax rows (nrb * rmax >= nrows).
std::vector<uint32_t> rct_row_blocks(const axspmv_csr *a, uint32_t nrb, uint32_t rmax) {
    const uint32_t n = a->nrows;
    std::vector<uint32_t> rbs(size_t(nrb) + 1, 0);
    for (uint32_t k = 1; k < nrb; ++k) {
        const uint64_t target = uint64_t(a->nnz) * k / nrb;
        uint64_t r = uint64_t(std::lower_bound(a->row, a->row + n + 1, target) - a->row);
        const uint64_t lo = std::max<uint64_t>(uint64_t(rbs[k - 1]) + 1,
                                               n > uint64_t(nrb - k) * rmax ? n - uint64_t(nrb - k) * rmax : 0);
        const uint64_t hi = std::min<uint64_t>(uint64_t(rbs[k - 1]) + rmax, n - (nrb - k));
        rbs[k] = uint32_t(std::min(std::max(r, lo), hi));
    }
    rbs[nrb] = n;
    return rbs;
}

// True when the rows, cut into blocks of `rows`, read narrow column bands:
// at most 1/8 of the nonzeros lie in blocks whose column span (max - min + 1)
// exceeds `cols`.  Such a matrix (banded, a stencil in natural order) keeps
// each block's x within one L2 already, and column slices would only leave
// most tiles empty.
bool narrow_bands(const axspmv_csr *a, uint32_t rows, uint64_t cols) {
    uint64_t wide = 0;
    for (uint64_t r0 = 0; r0 < a->nrows; r0 += rows) {
        const uint32_t r1 = uint32_t(std::min<uint64_t>(a->nrows, r0 + rows));
        const uint32_t b = a->row[r0], e = a->row[r1];
        if (b == e) continue;
        const auto mm = std::minmax_element(a->col + b, a->col + e);
        if (uint64_t(*mm.second) - *mm.first + 1 > cols) wide += e - b;
    }
    return wide * 8 <= a->nnz;
}

// RCT: row blocks of at most tile_rows rows and about equal nonzeros, column
// slices, entries of a tile sorted by (column, row).
int plan_rct(Plan &p, const axspmv_csr *a, hipStream_t s) {
    if (p.o.block > 1024 || p.o.block % 64)
        return fail(AXSPMV_E_INVALID, "RCT: block %u is not a multiple of 64 up to 1024", p.o.block);
    hipDeviceProp_t prop;
    size_t lds_max = 65536;
    uint64_t cus = 256;
    if (hipGetDeviceProperties(&prop, p.device) == hipSuccess) {
        if (prop.sharedMemPerBlock > 0) lds_max = prop.sharedMemPerBlock;
        if (prop.multiProcessorCount > 0) cus = uint64_t(prop.multiProcessorCount);
    }
    const uint32_t lds_rows = uint32_t(std::min<size_t>(65536, lds_max / p.lds_per_row));
    p.lds_max = lds_max;
    const uint32_t n = a->nrows;
    // Default tile height: as tall as the LDS allows (fewer row blocks, fewer
    // passes over x), but short enough that the tiles cover the 256 CUs.
    auto auto_rows = [&](uint32_t slices) {
        const uint64_t fill = ceil_div(ceil_div(uint64_t(n) * slices, 256), 64) * 64;
        return uint32_t(std::min<uint64_t>(std::min<uint32_t>(16384, lds_rows), std::max<uint64_t>(64, fill)));
    };
    // Default slicing: one slice when x fits 2 MB (one XCD's L2 share) or
    // the row blocks read narrow column bands anyway; else see below.
    constexpr uint64_t slice_cols = (2ull << 20) / 8;
    uint32_t S = p.o.slices;
    bool banded = false;
    if (S == 0) {
        S = 1;
        if (a->ncols > slice_cols) {
            banded = narrow_bands(a, p.o.tile_rows ? p.o.tile_rows : auto_rows(1), slice_cols);
            if (!banded) {
                uint32_t s2 = 1;  // slices of at most 2 MB of x
                while (s2 < 8 && ceil_div(a->ncols, s2) > slice_cols) s2 *= 2;
                if (p.lds_per_row > 8) {
                    // The plan of A^T behind a fused value gradient holds 16 B
                    // per tile row; 8192-row tiles with half the slices measured
                    // fastest for it (forward + adjoint on the 2^20-row
                    // power-law matrix 241 us against 269 us at 10240 rows and 4
                    // slices; 702 against 745 us on the permuted 27-point
                    // stencil on 128^3), and for the adjoint alone.
                    S = std::max(1u, s2 / 2);
                } else {
                    // Every slice adds each row's sum into y with an FP64
                    // atomic, so take the fewest slices that still give every
                    // CU a tile of full height (row blocks x slices >= CUs),
                    // but no more than the 2 MB rule asks and no fewer than
                    // slices of 8 MiB of x.  The permuted 27-point stencil on
                    // 128^3 (128 full-height row blocks): 2 slices 229.6 us,
                    // 4 slices 239.4 us, 8 slices 269.5 us; the 2^20-row
                    // power-law matrix (64 row blocks) keeps 4 slices (2: 121 us,
                    // half the CUs idle; 4: 74 us).
                    uint32_t s8 = 1;
                    while (s8 < 8 && ceil_div(a->ncols, s8) > (8ull << 20) / 8) s8 *= 2;
                    const uint64_t nrb_full = ceil_div(n, p.o.tile_rows ? p.o.tile_rows : std::min<uint32_t>(16384, lds_rows));
                    uint32_t sf = 1;
                    while (sf < 8 && nrb_full * sf < cus) sf *= 2;
                    S = std::max(s8, std::min(sf, s2));
                }
            }
        }
    }
    if (S != 1 && S != 2 && S != 4 && S != 8) return fail(AXSPMV_E_INVALID, "RCT: slices must be 1, 2, 4 or 8");
    // Narrow bands reuse x within a band whatever the tile height: 512-row
    // tiles (4 KB of LDS) let many workgroups share a CU and even out the
    // end of the launch, which pays on matrices mixing short and long rows
    // (256 banded matrices as one block-diagonal plan, row lengths varying
    // by 60 % of their mean, 16-bit keys: 985 us at 512 rows, 973 at 256,
    // 999 at 1024, 1035 at 2048; 32-bit keys: 1170 us at 512, 1227 at 2048,
    // 1390 at 16384).  Rows of one length (a mesh: the 128^3 stencil
    // renumbered by reverse Cuthill-McKee, 7 %) take the tallest tiles that
    // still cover the CUs: 168 against 180 us at 2048 rows, with the
    // non-temporal stream loads below.
    uint32_t Rt = p.o.tile_rows ? p.o.tile_rows : auto_rows(S);
    if (!p.o.tile_rows && banded) {
        double sum = 0.0, sq = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum, sq)
        for (uint32_t i = 0; i < n; ++i) {
            const double l = double(a->row[i + 1] - a->row[i]);
            sum += l;
            sq += l * l;
        }
        const double mean = n ? sum / n : 0.0, var = n ? std::max(0.0, sq / n - mean * mean) : 0.0;
        if (!(mean > 0.0 && std::sqrt(var) < 0.25 * mean)) Rt = std::min<uint32_t>(512, auto_rows(1));
    }
    if (p.lds_per_row > 8)  // a transposed plan's tiles are cut to fit (8192 rows when automatic)
        Rt = std::min(Rt, p.o.tile_rows || banded ? lds_rows : std::min<uint32_t>(lds_rows, 8192));
    if (Rt == 0 || Rt > lds_rows)
        return fail(AXSPMV_E_UNSUPPORTED, "RCT: %u rows per tile do not fit %zu bytes of LDS", Rt, lds_max);
    const uint32_t nrb = uint32_t(ceil_div(n, Rt));
    // Blocks keep at most Rt rows (so the packed key keeps its bits); the
    // balancing uses whatever slack ceil(n / Rt) * Rt - n leaves.
    const std::vector<uint32_t> rbs = rct_row_blocks(a, nrb, Rt);
    uint32_t R = 1;  // rows of the largest block: the LDS accumulator size
    for (uint32_t k = 0; k < nrb; ++k) R = std::max(R, rbs[k + 1] - rbs[k]);
    const uint32_t C = uint32_t(std::max<uint64_t>(1, ceil_div(a->ncols, S)));
    const uint64_t ntile = uint64_t(nrb) * S;
    // count entries per tile
    std::vector<uint64_t> cnt(ntile + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t rb = 0; rb < nrb; ++rb)
        for (uint32_t j = a->row[rbs[rb]]; j < a->row[rbs[rb + 1]]; ++j) cnt[uint64_t(rb) * S + a->col[j] / C + 1]++;
    for (uint64_t t = 0; t < ntile; ++t) cnt[t + 1] += cnt[t];
    const uint64_t nnz = a->nnz;
    std::vector<uint32_t> col(nnz ? nnz : 1);
    std::vector<uint16_t> lrow(nnz ? nnz : 1);
    std::vector<double> val(nnz ? nnz : 1);
    // the value gradient comes out in entry order: keep each entry's CSR position
    const bool keep_pos = p.keep_pos;
    std::vector<uint32_t> pos(keep_pos && nnz ? nnz : 1);
#pragma omp parallel
    {
        std::vector<std::pair<uint64_t, uint32_t>> key;  // (col << 16 | lrow, CSR position)
#pragma omp for schedule(dynamic, 1)
        for (uint32_t rb = 0; rb < nrb; ++rb) {
            const uint32_t r0 = rbs[rb], r1 = rbs[rb + 1];
            key.clear();
            for (uint32_t r = r0; r < r1; ++r)
                for (uint32_t j = a->row[r]; j < a->row[r + 1]; ++j)
                    key.emplace_back((uint64_t(a->col[j]) << 16) | (r - r0), j);
            std::sort(key.begin(), key.end());
            // sorted by column, so the slices come out in order
            uint64_t q = cnt[uint64_t(rb) * S];
            for (const auto &k : key) {
                col[q] = uint32_t(k.first >> 16);
                lrow[q] = uint16_t(k.first & 0xffff);
                val[q] = a->val[k.second];
                if (keep_pos) pos[q] = k.second;
                ++q;
            }
        }
    }
    // the kernel indexes entries in 32 bits with up to 2 * 8 * 1024 past the end
    if (cnt[ntile] >= (1ull << 32) - (1u << 16)) return fail(AXSPMV_E_OVERFLOW, "RCT: too many entries");
    std::vector<uint32_t> tptr(ntile + 1);
    for (uint64_t t = 0; t <= ntile; ++t) tptr[t] = uint32_t(cnt[t]);
    const size_t vs = p.o.vtype == AXSPMV_F32 ? 4 : 8;
    // Pack (column offset, row within the tile) into one word when both fit
    // in 32 bits.  The offset is taken from a base column per run of 1024
    // entries, counting round from the last column to column 0: a run's
    // columns rise tile by tile, so they cover a short stretch of columns,
    // far shorter than a tile's.  The base is the run's smallest column, or,
    // where the run passes from the last slice of a row block into the first
    // of the next (its columns fall back towards 0), its first column.
    auto bits = [](uint64_t v) {  // bits to hold values 0 .. v-1
        uint32_t b = 1;
        while ((1ull << b) < v) ++b;
        return b;
    };
    constexpr uint32_t SEG_LOG = 10;  // entries per run: 1 << SEG_LOG
    const uint64_t nseg = std::max<uint64_t>(1, ceil_div(nnz + (nnz & 1), uint64_t(1) << SEG_LOG));
    std::vector<uint32_t> sbase(nseg, 0);
    const uint64_t nc = std::max<uint64_t>(1, a->ncols);
    uint64_t span = 1;
#pragma omp parallel for schedule(static) reduction(max : span)
    for (uint64_t g = 0; g < nseg; ++g) {
        const uint64_t q0 = g << SEG_LOG, q1 = std::min<uint64_t>(nnz, q0 + (uint64_t(1) << SEG_LOG));
        if (q0 >= nnz) continue;
        const auto mm = std::minmax_element(col.begin() + q0, col.begin() + q1);
        const uint64_t lin = uint64_t(*mm.second) - *mm.first + 1;  // from the smallest column
        uint64_t cyc = 1;                                            // round from the first column
        for (uint64_t q = q0; q < q1; ++q) cyc = std::max<uint64_t>(cyc, (col[q] + nc - col[q0]) % nc + 1);
        sbase[g] = lin <= cyc ? *mm.first : col[q0];
        span = std::max(span, std::min(lin, cyc));
    }
    const uint32_t rbits = bits(R), cbits = bits(span);
    const bool packed = rbits + cbits <= 32 && !(p.o.flags & AXSPMV_PLAN_RCT_SPLIT);
    // Narrow bands: one 16-bit key per entry, the column as an offset from
    // the smallest column of its run of 128 entries and, above it, the
    // entry's row minus its column less the smallest such difference, where
    // both fit 16 bits together -- 10 bytes per FP64 entry instead of 12.
    // Entries of a band lie within a few columns of the diagonal, and a run
    // of 128 of them covers few columns, also where it passes from one tile
    // to the next (its columns step back by about the band's width).
    constexpr uint32_t DSEG_LOG = 7;  // entries per run: 1 << DSEG_LOG
    const uint64_t nseg16 = std::max<uint64_t>(1, ceil_div(nnz + (nnz & 1), uint64_t(1) << DSEG_LOG));
    std::vector<uint32_t> sbase16;
    int64_t dlo = 0;
    uint32_t cbits16 = 0;
    bool diag = packed && nnz > 0 && !(p.o.flags & AXSPMV_PLAN_RCT_WIDE_KEY);
    if (diag) {
        int64_t lo = INT64_MAX, hi = INT64_MIN;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : lo) reduction(max : hi)
        for (uint32_t rb = 0; rb < nrb; ++rb)
            for (uint64_t q = cnt[uint64_t(rb) * S]; q < cnt[uint64_t(rb + 1) * S]; ++q) {
                const int64_t d = int64_t(rbs[rb]) + lrow[q] - int64_t(col[q]);
                lo = std::min(lo, d);
                hi = std::max(hi, d);
            }
        sbase16.assign(nseg16, 0);
        uint64_t span16 = 1;
#pragma omp parallel for schedule(static) reduction(max : span16)
        for (uint64_t g = 0; g < nseg16; ++g) {
            const uint64_t q0 = g << DSEG_LOG, q1 = std::min<uint64_t>(nnz, q0 + (uint64_t(1) << DSEG_LOG));
            if (q0 >= nnz) continue;
            const auto mm = std::minmax_element(col.begin() + q0, col.begin() + q1);
            sbase16[g] = *mm.first;
            span16 = std::max<uint64_t>(span16, uint64_t(*mm.second) - *mm.first + 1);
        }
        dlo = lo;
        cbits16 = bits(span16);
        diag = cbits16 + bits(uint64_t(hi - lo) + 1) <= 16;
    }
    const uint64_t kbytes = diag ? 2 : packed ? 4 : 6;  // column and row per entry
    // one spare entry when nnz is odd: a kernel may load entries in aligned pairs
    const uint64_t nnz_pad = nnz + (nnz & 1);
    void *dv = p.dalloc(vs * nnz_pad);
    auto *dt = static_cast<uint32_t *>(p.dalloc(4 * tptr.size()));
    auto *drb = static_cast<uint32_t *>(p.dalloc(4 * rbs.size()));
    if (!dv || !dt || !drb) return AXSPMV_E_NOMEM;
    if (keep_pos) {
        pos.resize(nnz);
        p.rct_pos = std::move(pos);
    }
    AX_TRY(p.upload(drb, rbs.data(), 4 * rbs.size(), s));
    if (p.o.vtype == AXSPMV_F32) {
        std::vector<float> v(val.begin(), val.begin() + nnz);
        AX_TRY(p.upload(dv, v.data(), 4 * nnz, s));
    } else {
        AX_TRY(p.upload(dv, val.data(), 8 * nnz, s));
    }
    AX_TRY(p.upload(dt, tptr.data(), 4 * tptr.size(), s));
    uint32_t *dc = nullptr, *dk = nullptr, *dcb = nullptr;
    uint16_t *dl = nullptr, *dk16 = nullptr;
    if (diag) {
        std::vector<uint16_t> key(nnz);
#pragma omp parallel for schedule(dynamic, 1)
        for (uint32_t rb = 0; rb < nrb; ++rb)
            for (uint64_t q = cnt[uint64_t(rb) * S]; q < cnt[uint64_t(rb + 1) * S]; ++q) {
                const uint64_t d = uint64_t(int64_t(rbs[rb]) + lrow[q] - int64_t(col[q]) - dlo);
                key[q] = uint16_t(d << cbits16 | (col[q] - sbase16[q >> DSEG_LOG]));
            }
        dk16 = static_cast<uint16_t *>(p.dalloc(2 * nnz_pad));
        dcb = static_cast<uint32_t *>(p.dalloc(4 * nseg16));
        if (!dk16 || !dcb) return AXSPMV_E_NOMEM;
        AX_TRY(p.upload(dk16, key.data(), 2 * nnz, s));
        AX_TRY(p.upload(dcb, sbase16.data(), 4 * nseg16, s));
    } else if (packed) {
        std::vector<uint32_t> key(nnz ? nnz : 1);
#pragma omp parallel for schedule(static)
        for (uint64_t q = 0; q < nnz; ++q)
            key[q] = uint32_t(((col[q] + nc - sbase[q >> SEG_LOG]) % nc) << rbits) | lrow[q];
        dk = static_cast<uint32_t *>(p.dalloc(4 * nnz_pad));
        dcb = static_cast<uint32_t *>(p.dalloc(4 * nseg));
        if (!dk || !dcb) return AXSPMV_E_NOMEM;
        AX_TRY(p.upload(dk, key.data(), 4 * nnz, s));
        AX_TRY(p.upload(dcb, sbase.data(), 4 * nseg, s));
    } else {
        dc = static_cast<uint32_t *>(p.dalloc(4 * nnz_pad));
        dl = static_cast<uint16_t *>(p.dalloc(2 * nnz_pad));
        if (!dc || !dl) return AXSPMV_E_NOMEM;
        AX_TRY(p.upload(dc, col.data(), 4 * nnz, s));
        AX_TRY(p.upload(dl, lrow.data(), 2 * nnz, s));
    }
    uint32_t block = p.o.block;
    // 512 threads (8 waves, one workgroup per CU with R = 16384) measured
    // fastest on the 2^20-row power-law matrix (82K entries per tile; 1024 is
    // 2-3 % slower); tiles of few entries want fewer lanes (graph-timed: the
    // Laplacian on 256^2, 1.3K per tile, 3.69 -> 3.05 us at 128; a 2^16-row
    // power-law matrix, 5.1K per tile, 9.17 -> 7.83 us at 256).
    if (block == 0) {
        const uint64_t per_tile = nnz / std::max<uint64_t>(1, ntile);
        block = per_tile < 2048 ? 128 : per_tile < 16384 ? 256 : 512;
    }

    // S > 1: every tile writes its rows' sums when its loop ends, and all
    // tiles end together.  Where that write-out is heavy against the loop
    // (S n row sums >= nnz / 6), storing them in slice partials and merging
    // those in a second kernel beats adding them into a cleared y with FP64
    // atomics; where it is light, the merge's extra pass over S + 1 vectors
    // costs more than the adds.  Measured on one box, whole SpMV: the 2^20-row
    // power-law matrix (nnz 21 M) with 4 slices 60 against 66 us, 8 slices
    // 74 against 88 us, 2 slices 87 against 83 us; the permuted 27-point
    // stencil on 128^3 (nnz 56 M) with 2 slices 231 against 224 us, 4 slices
    // 242 against 240 us, 8 slices 267 against 274 us (partials against
    // atomics).  The plan of A^T, which adds into x_bar from its fused
    // value-gradient sweep, keeps the atomic adds.
    double *dpart = nullptr;
    if (S > 1 && p.lds_per_row == 8 && !(p.o.flags & AXSPMV_PLAN_RCT_ATOMIC) && 6ull * S * n >= nnz) {
        dpart = static_cast<double *>(p.dalloc(8ull * S * n));
        if (!dpart) return AXSPMV_E_NOMEM;
    }
    p.rct = launch::RctArgs{dv, dc, dl, dk, dcb, dt, drb, n, R, S, nrb, p.o.vtype, block, rbits, a->ncols, dpart};
    // A matrix stream larger than the 256 MB Infinity Cache comes from HBM
    // on every SpMV whatever the loads do: there tall tiles (one workgroup
    // per CU) stream it with non-temporal loads (the 128^3 stencil
    // renumbered by reverse Cuthill-McKee, 8192-row tiles: 184 -> 168 us per
    // SpMV; the permuted stencil's 16384-row tiles: unchanged).  A smaller
    // stream stays cached between SpMVs, and non-temporal loads give that up
    // (the 2^20-row power-law matrix, 252 MB: 58 -> 89 us); short tiles,
    // several workgroups per CU, measured 1 % slower with them (256 banded
    // matrices, 2048-row tiles: 1236 -> 1249 us).
    p.rct.nt_stream = (vs + kbytes) * nnz > (256ull << 20) && R >= 4096 ? 1 : 0;
    p.rct.key16 = dk16;
    p.rct.cbits16 = cbits16;
    p.rct.dmin = uint32_t(uint64_t(dlo));
    p.slots = nnz;
    // matrix stream; with S > 1 the partials written and read back and y
    // written by the merge, or the clear of y and S atomic adds per row
    p.stream_bytes = (vs + kbytes) * nnz + 4 * tptr.size() +
                     (S == 1 ? 0 : dpart ? 8ull * (2 * S + 1) * n : 8ull * (S + 1) * n);
    p.launches = S > 1 ? 2 : 1;
    p.nblocks = uint32_t(ntile);
    p.kind = Plan::RCT;
    std::snprintf(p.name, sizeof p.name, "rct_r%u_s%u%s%s", Rt, S, diag ? "_k16" : packed ? "" : "_split",
                  S > 1 && !dpart && p.lds_per_row == 8 ? "_atomic" : "");
    return AXSPMV_OK;
}

launch::SegArgs seg_args(const Plan &p, const uint32_t *rowsrc) {
    launch::SegArgs g{};
    g.vplane = p.d_vplane;
    g.cplane = p.d_cplane;
    g.rowsrc = rowsrc;
    g.blocks = p.d_blocks;
    g.nblocks = p.nblocks;
    g.empty = p.d_empty;
    g.nempty = p.nempty;
    g.thw = p.thw;
    g.th = p.th;
    g.vtype = p.o.vtype;
    g.block = p.o.block;
    g.win = p.d_win;
    g.nt = p.seg_nt;
    return g;
}

}  // namespace

extern "C" int axspmv_opts_default(uint32_t format, axspmv_opts *o) {
    if (!o) return fail(AXSPMV_E_INVALID, "opts_default: NULL");
    std::memset(o, 0, sizeof *o);
    o->format = format;
    o->kernel = AXSPMV_KER_DEFAULT;
    o->mode = AXSPMV_AXT_UNC;
    o->thw = 4;
    o->th = 1;
    o->bs = 256;
    o->hbs = format == AXSPMV_FMT_CSR || format == AXSPMV_FMT_AXC ? 0 : 64;  // picked from the mean row length
    o->chunk = 64;
    o->reorder = 1;
    o->vtype = AXSPMV_F64;
    o->block = format == AXSPMV_FMT_RCT ? 0 : 256;  // RCT: 512 for SpMV, 1024 for SpMM
    o->budget = 0;
    if (format > AXSPMV_FMT_RCT) return fail(AXSPMV_E_INVALID, "opts_default: unknown format %u", format);
    return AXSPMV_OK;
}

namespace {

int plan_create(const axspmv_csr *a, const axspmv_opts *o, void *stream, bool transposed, axspmv_plan **out);

// Mean |row - column| over the entries.
double mean_diagonal_distance(const axspmv_csr *a) {
    double sum = 0.0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : sum)
    for (uint32_t i = 0; i < a->nrows; ++i)
        for (uint32_t j = a->row[i]; j < a->row[i + 1]; ++j) sum += a->col[j] > i ? a->col[j] - i : i - a->col[j];
    return a->nnz ? sum / a->nnz : 0.0;
}

// dst[k] = src[perm[k]] in two passes of launch::permute_chunks through a
// staging vector t, in chunks of L = 8192 positions (64 KB of LDS).  t holds
// the values chunk of destinations by chunk (t[c L ..] feeds dst[c L ..]),
// inside a chunk by source block, then by destination.  Pass 1, workgroup s
// = source block s (src[s L ..]): reads the block, writes each value to its
// place in t (w1), the block's values for one destination chunk side by
// side (b1 = the value's offset in the block, in the order of w1).  Pass 2,
// workgroup c: reads t[c L ..] whole, dst[k] = that chunk's value b2[k].
// Measured on the permuted 128^3 stencil (2M entries, about 32 values per
// run; kernel trace): 11.3 + 8.3 us per vector, against 34 us for one plain
// gather (each 8-byte read pulls a whole line from the Infinity Cache), 29
// us for two XCD-segmented gathers (each XCD reading an L2-resident eighth
// of the source at random), 20.5 us with the runs on the reading side and
// 23.8 us with chunks of 4096.
int upload_perm(Plan &p, const std::vector<uint32_t> &perm, Plan::Perm &out, hipStream_t s) {
    const uint32_t n = uint32_t(perm.size());
    const uint32_t L = std::max(1u, std::min<uint32_t>(8192, n));
    const uint64_t nch = ceil_div(n, L);
    // t position of every destination: chunk by chunk, sorted by (source block, k)
    std::vector<uint32_t> tp(n ? n : 1);
    std::vector<uint16_t> b1(n ? n : 1), b2(n ? n : 1);
#pragma omp parallel for schedule(static)
    for (uint64_t c = 0; c < nch; ++c) {
        const uint32_t k0 = uint32_t(c * L), k1 = uint32_t(std::min<uint64_t>(n, k0 + uint64_t(L)));
        std::vector<std::pair<uint32_t, uint32_t>> v;  // (source block, destination)
        v.reserve(k1 - k0);
        for (uint32_t k = k0; k < k1; ++k) v.emplace_back(perm[k] / L, k);
        std::sort(v.begin(), v.end());
        for (uint32_t i = 0; i < k1 - k0; ++i) {
            tp[v[i].second] = k0 + i;
            b2[v[i].second] = uint16_t(i);
        }
    }
    // pass 1: each source block's values in order of their t positions (a
    // counting pass over t: destination k sits at t position tp[k])
    std::vector<uint32_t> kat(n ? n : 1), w1(n ? n : 1), next(nch);
    for (uint32_t k = 0; k < n; ++k) kat[tp[k]] = k;
    for (uint64_t c = 0; c < nch; ++c) next[c] = uint32_t(c * L);
    for (uint32_t t = 0; t < n; ++t) {
        const uint32_t k = kat[t], blk = perm[k] / L, j = next[blk]++;
        w1[j] = t;
        b1[j] = uint16_t(perm[k] - blk * L);
    }
    out.L = L;
    out.b1 = static_cast<uint16_t *>(p.dalloc(2ull * n));
    out.b2 = static_cast<uint16_t *>(p.dalloc(2ull * n));
    out.w1 = static_cast<uint32_t *>(p.dalloc(4ull * n));
    if (!out.b1 || !out.b2 || !out.w1) return AXSPMV_E_NOMEM;
    AX_TRY(p.upload(out.b1, b1.data(), 2ull * n, s));
    AX_TRY(p.upload(out.b2, b2.data(), 2ull * n, s));
    return p.upload(out.w1, w1.data(), 4ull * n, s);
}

// AXSPMV_PLAN_REORDER: Q = the reverse Cuthill-McKee order of A, the plan of
// B = Q A Q^T (every other option as given) in `inner`, the permutation and
// its inverse on the device, and the vectors the gathers go through.
int plan_reordered(const axspmv_csr *a, const axspmv_opts *o, void *stream, axspmv_plan **out) {
    if (a->nrows != a->ncols)
        return fail(AXSPMV_E_INVALID, "AXSPMV_PLAN_REORDER: the %u x %u matrix is not square", a->nrows, a->ncols);
    std::vector<uint32_t> q, pos;
    AX_TRY(rcm_order(a, q));
    axspmv_csr *b = nullptr;
    AX_TRY(csr_permute_sym(a, q, &b, &pos));
    axspmv_opts oi = *o;
    oi.flags &= ~AXSPMV_PLAN_REORDER;
    // Worth it only where the new order brings the entries at least twice
    // as close to the diagonal on average (a random matrix gains nothing, a
    // banded one is already close): else the plan of A itself.
    if (2.0 * mean_diagonal_distance(b) > mean_diagonal_distance(a)) {
        axspmv_csr_free(b);
        return plan_create(a, &oi, stream, false, out);
    }
    axspmv_plan *pi = nullptr;
    const int rc = plan_create(b, &oi, stream, false, &pi);
    axspmv_csr_free(b);
    if (rc) return rc;
    auto p = std::make_unique<Plan>();
    p->inner = pi;
    p->kind = Plan::REORDERED;
    p->o = pi->o;
    p->o.flags |= AXSPMV_PLAN_REORDER;
    p->device = pi->device;
    p->nrows = a->nrows;
    p->ncols = a->ncols;
    p->nnz = a->nnz;
    p->thw = pi->thw;
    p->th = pi->th;
    const uint32_t n = a->nrows;
    std::vector<uint32_t> qi(n);
    for (uint32_t k = 0; k < n; ++k) qi[q[k]] = k;
    hipStream_t s = static_cast<hipStream_t>(stream);
    AX_TRY(upload_perm(*p, q, p->to_b, s));
    AX_TRY(upload_perm(*p, qi, p->from_b, s));
    p->d_perm = static_cast<uint32_t *>(p->dalloc(4ull * n));
    p->d_iperm = static_cast<uint32_t *>(p->dalloc(4ull * n));
    if (!p->d_perm || !p->d_iperm) return AXSPMV_E_NOMEM;
    AX_TRY(p->upload(p->d_perm, q.data(), 4ull * n, s));
    AX_TRY(p->upload(p->d_iperm, qi.data(), 4ull * n, s));
    p->nscratch = pi->T ? 6 : 3;
    for (uint32_t k = 0; k < p->nscratch; ++k) {
        p->d_scr[k] = static_cast<double *>(p->dalloc(8ull * n));
        if (!p->d_scr[k]) return AXSPMV_E_NOMEM;
    }
    p->bpos = std::move(pos);
    std::snprintf(p->name, sizeof p->name, "%.59s_rcm", pi->name);
    p->slots = pi->slots;
    p->beta = pi->beta;
    p->device_bytes += pi->device_bytes;
    p->alg_bytes = pi->alg_bytes;
    // the inner plan's bytes, then per vector permuted (x into B's order
    // before the SpMV, y back after it) two passes of 8 B read and 8 B
    // written, with 2 + 6 B of indices.  (RCT tiles storing their rows straight into A's
    // numbering, y[q[k]], measured slower on the permuted 128^3 stencil: the
    // tile kernel took 38 us more with ordinary scattered stores, 56 us more
    // with non-temporal ones, against 30 us for the two passes.)
    p->stream_bytes = pi->stream_bytes + 2ull * 40 * n;  // 2 vectors x (16 + 16 + 8) B
    p->launches = pi->launches + 4;
    p->nblocks = pi->nblocks;
    p->staged_blocks = pi->staged_blocks;
    p->rowwise = pi->rowwise;
    *out = p.release();
    return AXSPMV_OK;
}

// axspmv_plan_create; `transposed` marks the plan of A^T that a plan with
// AXSPMV_PLAN_TRANSPOSE builds (RCT: entry positions kept, 16 B of LDS per
// tile row for the fused value gradient).
int plan_create(const axspmv_csr *a, const axspmv_opts *o, void *stream, bool transposed, axspmv_plan **out) {
    if (!a || !o || !out) return fail(AXSPMV_E_INVALID, "plan_create: NULL argument");
    *out = nullptr;
    // Every kernel gathers x[col] without a bounds check: offsets and
    // columns are checked before anything is built.
    AX_TRY(check_csr_bounds(a, "plan_create"));
    if (o->vtype > AXSPMV_F32) return fail(AXSPMV_E_INVALID, "plan_create: unknown value type %u", o->vtype);
    // A group of lanes per row (CSR vector, AXC): the launch counts nrows x
    // lanes work-items, which must fit the 32-bit grid size of a dispatch.
    if ((o->format == AXSPMV_FMT_CSR && o->kernel == AXSPMV_KER_STREAM) ||
        (o->format == AXSPMV_FMT_AXC && !(o->flags & AXSPMV_PLAN_AX_LAYOUT))) {
        const uint64_t g = o->hbs ? o->hbs : lanes_for_mean(a);
        if (uint64_t(a->nrows) * g >= (1ull << 32))
            return fail(AXSPMV_E_OVERFLOW, "plan_create: %u rows x %llu lanes per row exceed a 32-bit grid", a->nrows,
                        (unsigned long long)g);
    }
    if (o->flags & AXSPMV_PLAN_REORDER) return plan_reordered(a, o, stream, out);
    auto p = std::make_unique<Plan>();
    p->o = *o;
    if (transposed) {
        p->keep_pos = true;
        p->lds_per_row = 16;
    }
    if (p->o.th == 0) p->o.th = 1;
    if (p->o.thw == 0) p->o.thw = p->o.th == 1 ? 4 : 64;
    if (p->o.hbs == 0 && p->o.format != AXSPMV_FMT_CSR && p->o.format != AXSPMV_FMT_AXC) p->o.hbs = 64;
    if (p->o.chunk == 0) p->o.chunk = 64;
    if (p->o.bs == 0) p->o.bs = 256;
    if (hipGetDevice(&p->device) != hipSuccess) return fail(AXSPMV_E_HIP, "plan_create: no HIP device");
    p->nrows = a->nrows;
    p->ncols = a->ncols;
    p->nnz = a->nnz;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = AXSPMV_OK;
    switch (o->format) {
        case AXSPMV_FMT_CSR: rc = plan_csr(*p, a, s); break;
        case AXSPMV_FMT_K1: rc = plan_k1(*p, a, s); break;
        case AXSPMV_FMT_AXC: rc = plan_axc(*p, a, s); break;
        case AXSPMV_FMT_AXT: rc = plan_axt(*p, a, s); break;
        case AXSPMV_FMT_RCT: rc = plan_rct(*p, a, s); break;
        default: rc = fail(AXSPMV_E_INVALID, "plan_create: unknown format %u", o->format);
    }
    if (rc) return rc;
    if (p->o.flags & AXSPMV_PLAN_TRANSPOSE) AX_TRY(plan_add_transpose(p.get(), a, s));
    const size_t vs = p->o.vtype == AXSPMV_F32 ? 4 : 8;
    p->alg_bytes = (vs + 4) * uint64_t(a->nnz) + 4ull * (a->nrows + 1) + 8ull * a->ncols + 8ull * a->nrows;
    p->stream_bytes += 8ull * a->ncols + 8ull * a->nrows;
    p->beta = p->slots ? double(a->nnz) / double(p->slots) : 0.0;
    *out = p.release();
    return AXSPMV_OK;
}

}  // namespace

extern "C" int axspmv_plan_create(const axspmv_csr *a, const axspmv_opts *o, void *stream, axspmv_plan **out) {
    return plan_create(a, o, stream, false, out);
}

extern "C" int axspmv_plan_destroy(axspmv_plan *p) {
    delete p;
    return AXSPMV_OK;
}

extern "C" int axspmv_plan_get_info(const axspmv_plan *p, axspmv_plan_info *info) {
    if (!p || !info) return fail(AXSPMV_E_INVALID, "plan_get_info: NULL argument");
    std::memset(info, 0, sizeof *info);
    std::snprintf(info->name, sizeof info->name, "%s", p->name);
    info->nrows = p->nrows;
    info->ncols = p->ncols;
    info->nnz = p->nnz;
    info->format = p->o.format;
    info->slots = p->slots;
    info->device_bytes = p->device_bytes;
    info->stream_bytes = p->stream_bytes;
    info->alg_bytes = p->alg_bytes;
    info->beta = p->beta;
    info->launches = p->launches;
    info->nblocks = p->nblocks;
    info->staged_blocks = p->staged_blocks;
    return AXSPMV_OK;
}

namespace axspmv {

// dst = src permuted by q, staged in the plan's last scratch vector
static hipError_t permute(const Plan *p, const Plan::Perm &q, const double *src, double *dst, hipStream_t s) {
    double *t = p->d_scr[p->nscratch - 1];
    hipError_t e = launch::permute_chunks(q.w1, q.b1, p->nrows, q.L, src, t, s);
    if (e == hipSuccess) e = launch::permute_chunks(nullptr, q.b2, p->nrows, q.L, t, dst, s);
    return e;
}

hipError_t plan_launch(axspmv_plan *p, const double *x, double *y, hipStream_t s) {
    switch (p->kind) {
        case Plan::CSR_NAIVE:
            return launch::csr_naive(p->nrows, p->d_val, p->o.vtype, p->d_col, p->d_row, x, y, p->o.block, s);
        case Plan::CSR_VECTOR:
            return launch::csr_vector(p->nrows, p->d_val, p->o.vtype, p->d_col, p->d_row, p->lanes, p->spmm.long_slots,
                                      p->spmm.longrows, p->spmm.nlong, x, y, s, p->seg_nt);
        case Plan::AXT_UNC_H1: return launch::axt_unc_h1(seg_args(*p, p->d_sec), x, y, s);
        case Plan::AXT_UNC_H1_SLICED: {
            launch::SegArgs g = seg_args(*p, p->d_sec);
            g.nrows_add = p->nrows;
            g.mid = p->mid;
            g.part = p->d_part;
            g.S = p->slices;
            return launch::axt_unc_h1(g, x, y, s);
        }
        case Plan::AXT_UNC_H: return launch::axt_unc_h(seg_args(*p, p->d_sec), x, y, s);
        case Plan::AXT_COM: return launch::axt_com(p->com, x, y, s);
        case Plan::K1: return launch::k1(p->k1, x, y, s);
        case Plan::AXC: return launch::axc(p->axc, x, y, s);
        case Plan::RCT: return launch::rct(p->rct, x, y, s, p->mid);
        case Plan::REORDERED: {
            hipError_t e = permute(p, p->to_b, x, p->d_scr[0], s);
            if (e == hipSuccess) e = plan_launch(p->inner, p->d_scr[0], p->d_scr[1], s);
            if (e == hipSuccess) e = permute(p, p->from_b, p->d_scr[1], y, s);
            return e;
        }
        case Plan::AXT_UNC_H1_AX: {
            const hipError_t e = launch::x_refresh(p->d_axmap, p->slots, ilog2(p->thw), p->ncols, x, p->d_ax, s);
            if (e != hipSuccess) return e;
            launch::SegArgs g = seg_args(*p, p->d_sec);
            g.vplane = p->d_ax;
            g.win = nullptr;
            return launch::axt_unc_h1_ax(g, y, s);
        }
        default: return hipErrorInvalidValue;
    }
}

}  // namespace axspmv

extern "C" int axspmv_spmv(axspmv_plan *p, const double *d_x, double *d_y, void *stream) {
    if (!p || (!d_x && p->ncols) || (!d_y && p->nrows)) return fail(AXSPMV_E_INVALID, "spmv: NULL argument");
    const hipError_t e = plan_launch(p, d_x, d_y, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "spmv launch");
}

// ---------------------------------------------------------------- adjoint
namespace {

launch::SegArgs fused_args(const Plan &p) {
    launch::SegArgs g = seg_args(p, p.d_sec);
    g.blocks = p.d_fblocks;
    g.nblocks = p.nfblocks;
    g.win = nullptr;  // windows belong to the SpMV blocks
    return g;
}

hipError_t launch_vbar(Plan *p, const double *ybar, const double *x, double *vbar, hipStream_t s) {
    // RCT: over the plan of A^T (rows = columns of A), whose entry order is
    // the one the fused forward + adjoint writes: x rows in LDS, ybar gathered
    if (p->kind == Plan::RCT) return launch::rct_vbar(p->T->rct, x, ybar, vbar, s);
    if (p->kind == Plan::AXT_UNC_H1)
        return launch::axt_unc_h1_vbar(seg_args(*p, p->d_sec), p->nunits, x, ybar, vbar,
                                       launch::VbarArgs{p->d_rowptr, p->d_first}, s);
    return launch::csr_vbar(p->nrows, p->d_col, p->d_row, x, ybar, vbar, s);
}

hipError_t launch_fwd_adj(Plan *p, const double *x, const double *b, double *y, double *r, double *xbar,
                          double *vbar, bool fused, hipStream_t s) {
    hipError_t e;
    if (fused && p->kind != Plan::RCT) {  // RCT: r needs every slice's sums first
        if (p->kind == Plan::AXT_UNC_H1)
            e = launch::axt_unc_h1_fused(fused_args(*p), x, b, y, r, vbar, launch::VbarArgs{p->d_rowptr, p->d_first}, s);
        else
            e = launch::csr_fused(p->nrows, p->d_val, p->o.vtype, p->d_col, p->d_row, x, b, y, r, vbar, s);
        if (e != hipSuccess) return e;
        return plan_launch(p->T, r, xbar, s);
    }
    if (p->kind == Plan::RCT) {
        // r = y - b where y's rows are written (the slice merge, or the tile
        // kernel of one slice).  With the non-temporal v_bar stores of the
        // sweep over A^T, forward + adjoint on the 2^20-row power-law matrix,
        // one box: 232.6 -> 223.4 us.
        e = launch::rct(p->rct, x, y, s, nullptr, nullptr, nullptr, b, r);
    } else {
        e = plan_launch(p, x, y, s);
        if (e == hipSuccess) e = launch::residual(p->nrows, y, b, r, s);
    }
    if (fused && p->kind == Plan::RCT) {
        // x_bar = A^T r and v_bar in one sweep over the plan of A^T
        if (e == hipSuccess) e = launch::rct(p->T->rct, r, xbar, s, nullptr, x, vbar);
        return e;
    }
    if (e == hipSuccess) e = plan_launch(p->T, r, xbar, s);
    if (e == hipSuccess) e = launch_vbar(p, r, x, vbar, s);
    return e;
}

// RCT tiles hold R x ktile sums in LDS: the default k-tile is also cut to
// what fits.
uint32_t pick_ktile(const Plan *p, uint32_t k) {
    if (p->o.ktile) return p->o.ktile;
    uint32_t kt = 8;
    while (kt < 64 && kt < k) kt *= 2;
    if (p->kind == Plan::RCT)
        while (kt > 8 && uint64_t(p->rct.R) * kt * 8 > p->lds_max) kt /= 2;
    return kt;
}

bool adjoint_kind(const Plan *p) {
    return p->kind == Plan::AXT_UNC_H1 || p->kind == Plan::CSR_NAIVE || p->kind == Plan::RCT;
}

}  // namespace

namespace axspmv {

int plan_add_transpose(Plan *p, const axspmv_csr *a, hipStream_t s) {
    axspmv_csr *t = nullptr;
    std::vector<uint32_t> perm;  // position in A^T -> position in A
    AX_TRY(csr_transpose_perm(a, &t, &perm));
    axspmv_opts o = p->o;
    o.flags &= ~AXSPMV_PLAN_TRANSPOSE;
    axspmv_plan *pt = nullptr;
    const int rc = plan_create(t, &o, s, true, &pt);
    if (rc) {
        axspmv_csr_free(t);
        return rc;
    }
    p->T = pt;
    if (pt->kind == Plan::RCT && (p->o.flags & AXSPMV_PLAN_DETERMINISTIC)) {
        // x_bar = A^T y_bar without atomics: the AXT layout of A^T in column
        // slices of at most 2 MB of y_bar (at most 8), slice partials summed
        // in slice order.  Two slots per tile: on the column-uniform rows of
        // a power-law matrix's transpose the least padding per slice.
        axspmv_opts od{};
        axspmv_opts_default(AXSPMV_FMT_AXT, &od);
        od.thw = 2;
        od.vtype = p->o.vtype;
        od.flags = AXSPMV_PLAN_DETERMINISTIC;
        uint32_t S = 1;
        while (S < 8 && 8ull * t->ncols > (2ull << 20) * S) S *= 2;
        od.slices = S;
        axspmv_plan *pd = nullptr;
        const int rd = plan_create(t, &od, s, false, &pd);
        if (rd) {
            axspmv_csr_free(t);
            return rd;
        }
        p->Tdet = pd;
    }
    axspmv_csr_free(t);
    if (pt->kind == Plan::RCT) {
        // v_bar comes out in the transposed plan's entry order: the position
        // in A of each of its entries
        p->rct_pos.resize(pt->rct_pos.size());
        for (size_t e = 0; e < pt->rct_pos.size(); ++e) p->rct_pos[e] = perm[pt->rct_pos[e]];
        std::vector<uint32_t>().swap(pt->rct_pos);
    }
    return AXSPMV_OK;
}

// A reordered plan: inputs gathered into B's numbering (x~ = Q x: x~[k] =
// x[perm[k]]), the operation on the plan of B, results gathered back (y =
// Q^T y~: y[i] = y~[iperm[i]]).  With B = Q A Q^T, A^T = Q^T B^T Q, so the
// adjoint takes the same two gathers; v_bar stays in the order of B's plan.
static hipError_t reordered_op(Plan *p, uint32_t op, const axspmv_operands &o, hipStream_t s) {
    double *const *w = p->d_scr;
    auto in = [&](const double *v, double *d) { return permute(p, p->to_b, v, d, s); };
    auto back = [&](const double *v, double *d) { return permute(p, p->from_b, v, d, s); };
    axspmv_operands io{};
    hipError_t e = hipSuccess;
    switch (op) {
        case AXSPMV_OP_SPMV: return plan_launch(p, o.x, o.y, s);
        case AXSPMV_OP_ADJOINT:
            io.ybar = w[0];
            io.xbar = w[1];
            e = in(o.ybar, w[0]);
            if (e == hipSuccess) e = plan_op(p->inner, op, io, s);
            if (e == hipSuccess) e = back(w[1], o.xbar);
            return e;
        case AXSPMV_OP_VBAR:
            io.ybar = w[0];
            io.x = w[1];
            io.vbar = o.vbar;
            e = in(o.ybar, w[0]);
            if (e == hipSuccess) e = in(o.x, w[1]);
            if (e == hipSuccess) e = plan_op(p->inner, op, io, s);
            return e;
        case AXSPMV_OP_FWD_ADJ:
        case AXSPMV_OP_FWD_ADJ_SPLIT:
            if (op == AXSPMV_OP_FWD_ADJ && p->inner->kind != Plan::RCT) {
                // the format's fused sweep takes b: x and b in, y, r and x_bar out
                io.x = w[0];
                io.b = w[1];
                io.y = w[2];
                io.r = w[3];
                io.xbar = w[4];
                io.vbar = o.vbar;
                e = in(o.x, w[0]);
                if (e == hipSuccess) e = in(o.b, w[1]);
                if (e == hipSuccess) e = plan_op(p->inner, op, io, s);
                if (e == hipSuccess) e = back(w[2], o.y);
                if (e == hipSuccess) e = back(w[3], o.r);
                if (e == hipSuccess) e = back(w[4], o.xbar);
                return e;
            }
            // four permutations instead of five: y comes back, r = y - b in
            // A's numbering, r goes in; then x_bar and v_bar from B's plan of
            // B^T (RCT: one sweep; otherwise the adjoint and the value
            // gradient), x_bar back.  y's way back and r's way in meet in
            // the same chunks of A's numbering: one kernel does the second
            // pass of the one, the residual and the first pass of the other
            // (y~'s vector, spent by then, takes r's runs).  With the
            // non-temporal v_bar stores of the RCT sweep, on one box: the
            // permuted 128^3 stencil, RCT, 483.4 -> 470.0 us.
            e = in(o.x, w[0]);
            if (e == hipSuccess) e = plan_launch(p->inner, w[0], w[1], s);
            if (p->from_b.L == p->to_b.L) {
                double *t = p->d_scr[p->nscratch - 1];
                const Plan::Perm &f = p->from_b, &g = p->to_b;
                if (e == hipSuccess) e = launch::permute_chunks(f.w1, f.b1, p->nrows, f.L, w[1], t, s);
                if (e == hipSuccess)
                    e = launch::back_residual_in(f.b2, g.w1, g.b1, p->nrows, f.L, t, o.b, o.y, o.r, w[1], s);
                if (e == hipSuccess) e = launch::permute_chunks(nullptr, g.b2, p->nrows, g.L, w[1], w[2], s);
            } else {
                if (e == hipSuccess) e = back(w[1], o.y);
                if (e == hipSuccess) e = launch::residual(p->nrows, o.y, o.b, o.r, s);
                if (e == hipSuccess) e = in(o.r, w[2]);
            }
            if (e != hipSuccess) return e;
            if (op == AXSPMV_OP_FWD_ADJ) {  // RCT: x_bar = B^T r~ and v_bar in one sweep over the plan of B^T
                e = launch::rct(p->inner->T->rct, w[2], w[3], s, nullptr, w[0], o.vbar);
            } else {
                io.ybar = w[2];
                io.xbar = w[3];
                e = plan_op(p->inner, AXSPMV_OP_ADJOINT, io, s);
                io.x = w[0];
                io.vbar = o.vbar;
                if (e == hipSuccess) e = plan_op(p->inner, AXSPMV_OP_VBAR, io, s);
            }
            if (e == hipSuccess) e = back(w[3], o.xbar);
            return e;
        case AXSPMV_OP_SPMM: {  // rows of X into B's order, Y's rows back
            const uint64_t nk = uint64_t(p->nrows) * o.k;
            if (!p->d_mm || p->mm_cap < 2 * nk) return hipErrorInvalidValue;  // plan_op_prepare first
            io.k = o.k;
            io.X = p->d_mm;
            io.Y = p->d_mm + nk;
            e = launch::gather_rows(p->d_perm, p->nrows, o.k, o.X, p->d_mm, s);
            if (e == hipSuccess) e = plan_op(p->inner, op, io, s);
            if (e == hipSuccess) e = launch::gather_rows(p->d_iperm, p->nrows, o.k, p->d_mm + nk, o.Y, s);
            return e;
        }
        case AXSPMV_OP_XREFRESH:
            io.x = w[0];
            e = in(o.x, w[0]);
            if (e == hipSuccess) e = plan_op(p->inner, op, io, s);
            return e;
        default: return hipErrorInvalidValue;
    }
}

hipError_t plan_op(Plan *p, uint32_t op, const axspmv_operands &o, hipStream_t s) {
    if (p->kind == Plan::REORDERED) return reordered_op(p, op, o, s);
    switch (op) {
        case AXSPMV_OP_SPMV: return plan_launch(p, o.x, o.y, s);
        case AXSPMV_OP_ADJOINT: return plan_launch(p->Tdet ? p->Tdet : p->T, o.ybar, o.xbar, s);
        case AXSPMV_OP_VBAR: return launch_vbar(p, o.ybar, o.x, o.vbar, s);
        case AXSPMV_OP_FWD_ADJ: return launch_fwd_adj(p, o.x, o.b, o.y, o.r, o.xbar, o.vbar, true, s);
        case AXSPMV_OP_FWD_ADJ_SPLIT: return launch_fwd_adj(p, o.x, o.b, o.y, o.r, o.xbar, o.vbar, false, s);
        case AXSPMV_OP_SPMM:
            if (p->kind == Plan::RCT) {
                // 1024-lane workgroups unless the plan sets the block: a tile
                // has ktile times the SpMV's LDS adds to hide
                launch::RctArgs a = p->rct;
                if (!p->o.block) a.block = 1024;
                return launch::rct_spmm(a, o.k, pick_ktile(p, o.k), o.X, o.Y, s);
            }
            return launch::spmm(p->spmm, o.k, pick_ktile(p, o.k), o.X, o.Y, s);
        case AXSPMV_OP_XREFRESH:
            return launch::x_refresh(p->d_axmap, p->slots, ilog2(p->thw), p->ncols, o.x, p->d_ax, s);
        default: return hipErrorInvalidValue;
    }
}

int plan_op_check(const Plan *p, uint32_t op, const axspmv_operands &o) {
    if (!p) return fail(AXSPMV_E_INVALID, "NULL plan");
    if (p->kind == Plan::REORDERED) {
        if (op > AXSPMV_OP_XREFRESH) return fail(AXSPMV_E_INVALID, "unknown operation %u", op);
        if (op == AXSPMV_OP_SPMM && uint64_t(p->nrows) * o.k >= (1ull << 40))
            return fail(AXSPMV_E_OVERFLOW, "SpMM: %u rows x k %u", p->nrows, o.k);
        if (op == AXSPMV_OP_SPMV && ((!o.x && p->ncols) || (!o.y && p->nrows)))
            return fail(AXSPMV_E_INVALID, "spmv: NULL argument");
        return plan_op_check(p->inner, op, o);  // the operation's own conditions, on the plan of B
    }
    const bool needT = op == AXSPMV_OP_ADJOINT || op == AXSPMV_OP_FWD_ADJ || op == AXSPMV_OP_FWD_ADJ_SPLIT ||
                       (op == AXSPMV_OP_VBAR && p->kind == Plan::RCT);
    if (needT && !p->T)
        return fail(AXSPMV_E_INVALID, "%s needs a plan created with AXSPMV_PLAN_TRANSPOSE",
                    op == AXSPMV_OP_VBAR ? "the RCT value gradient (it runs on the plan of A^T)" : "the adjoint");
    if ((op == AXSPMV_OP_VBAR || op == AXSPMV_OP_FWD_ADJ || op == AXSPMV_OP_FWD_ADJ_SPLIT) && !adjoint_kind(p))
        return fail(AXSPMV_E_UNSUPPORTED, "value gradient: plan %s is not CSR naive, AXT UNC th=1 or RCT", p->name);
    if (op == AXSPMV_OP_SPMM) {
        if (!p->rowwise && p->kind != Plan::RCT)
            return fail(AXSPMV_E_UNSUPPORTED, "SpMM: plan %s is not CSR, AXC, AXT UNC th=1 or RCT", p->name);
        if (o.k == 0) return fail(AXSPMV_E_INVALID, "SpMM: k must be > 0");
        const uint32_t kt = pick_ktile(p, o.k);
        if (kt != 8 && kt != 16 && kt != 32 && kt != 64) return fail(AXSPMV_E_INVALID, "SpMM: ktile %u not in 8/16/32/64", kt);
        if (p->kind != Plan::RCT && uint64_t(p->nrows) * kt >= (1ull << 32))
            return fail(AXSPMV_E_OVERFLOW, "SpMM: %u rows x ktile %u exceed a 32-bit grid", p->nrows, kt);
        if (p->kind == Plan::RCT && uint64_t(p->rct.R) * kt * 8 > p->lds_max)
            return fail(AXSPMV_E_UNSUPPORTED,
                        "SpMM: RCT tiles of %u rows x ktile %u need %llu bytes of LDS, %zu available (plan with "
                        "tile_rows <= %zu)",
                        p->rct.R, kt, (unsigned long long)(uint64_t(p->rct.R) * kt * 8), p->lds_max,
                        p->lds_max / (8 * size_t(kt)));
        if ((!o.X && p->ncols) || (!o.Y && p->nrows)) return fail(AXSPMV_E_INVALID, "SpMM: NULL operand");
    }
    if (op == AXSPMV_OP_XREFRESH) {
        if (p->kind != Plan::AXT_UNC_H1_AX)
            return fail(AXSPMV_E_UNSUPPORTED, "x refresh: plan %s was not created with AXSPMV_PLAN_AX_LAYOUT", p->name);
        if (!o.x && p->ncols) return fail(AXSPMV_E_INVALID, "x refresh: NULL x");
    }
    if (op > AXSPMV_OP_XREFRESH) return fail(AXSPMV_E_INVALID, "unknown operation %u", op);
    return AXSPMV_OK;
}

int plan_op_prepare(Plan *p, uint32_t op, const axspmv_operands &o) {
    if (p->kind != Plan::REORDERED || op != AXSPMV_OP_SPMM) return AXSPMV_OK;
    const uint64_t need = 2ull * p->nrows * o.k;
    if (p->d_mm && p->mm_cap >= need) return AXSPMV_OK;
    // A smaller buffer stays allocated until the plan goes: a graph captured
    // with it may still run.
    void *m = p->dalloc(8 * std::max<uint64_t>(need, 1));
    if (!m) return AXSPMV_E_NOMEM;
    p->d_mm = static_cast<double *>(m);
    p->mm_cap = need;
    return AXSPMV_OK;
}

}  // namespace axspmv

extern "C" int axspmv_adjoint(axspmv_plan *p, const double *d_ybar, double *d_xbar, void *stream) {
    axspmv_operands o{};
    o.ybar = d_ybar;
    o.xbar = d_xbar;
    AX_TRY(plan_op_check(p, AXSPMV_OP_ADJOINT, o));
    const hipError_t e = plan_op(p, AXSPMV_OP_ADJOINT, o, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "adjoint launch");
}

extern "C" int axspmv_vbar(axspmv_plan *p, const double *d_ybar, const double *d_x, double *d_vbar, void *stream) {
    axspmv_operands o{};
    o.ybar = d_ybar;
    o.x = d_x;
    o.vbar = d_vbar;
    AX_TRY(plan_op_check(p, AXSPMV_OP_VBAR, o));
    const hipError_t e = plan_op(p, AXSPMV_OP_VBAR, o, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "vbar launch");
}

extern "C" int axspmv_fwd_adj(axspmv_plan *p, const double *d_x, const double *d_b, double *d_y, double *d_r,
                              double *d_xbar, double *d_vbar, int fused, void *stream) {
    axspmv_operands o{};
    o.x = d_x;
    o.b = d_b;
    o.y = d_y;
    o.r = d_r;
    o.xbar = d_xbar;
    o.vbar = d_vbar;
    const uint32_t op = fused ? AXSPMV_OP_FWD_ADJ : AXSPMV_OP_FWD_ADJ_SPLIT;
    AX_TRY(plan_op_check(p, op, o));
    const hipError_t e = plan_op(p, op, o, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "fwd_adj launch");
}

extern "C" int axspmv_spmm(axspmv_plan *p, uint32_t k, const double *d_X, double *d_Y, void *stream) {
    axspmv_operands o{};
    o.k = k;
    o.X = d_X;
    o.Y = d_Y;
    AX_TRY(plan_op_check(p, AXSPMV_OP_SPMM, o));
    AX_TRY(plan_op_prepare(p, AXSPMV_OP_SPMM, o));
    const hipError_t e = plan_op(p, AXSPMV_OP_SPMM, o, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "spmm launch");
}

extern "C" int axspmv_plan_vbar_positions(const axspmv_plan *p, uint32_t *h_pos, uint64_t count) {
    if (!p || (count && !h_pos)) return fail(AXSPMV_E_INVALID, "vbar_positions: NULL argument");
    if (count != p->nnz) return fail(AXSPMV_E_INVALID, "vbar_positions: count %llu != nnz %u", (unsigned long long)count, p->nnz);
    if (p->kind == Plan::REORDERED) {  // positions in B, then in A
        AX_TRY(axspmv_plan_vbar_positions(p->inner, h_pos, count));
        for (uint64_t k = 0; k < count; ++k) h_pos[k] = p->bpos[h_pos[k]];
        return AXSPMV_OK;
    }
    if (p->kind == Plan::RCT) {
        if (!p->T)
            return fail(AXSPMV_E_UNSUPPORTED, "vbar_positions: RCT plan created without AXSPMV_PLAN_TRANSPOSE");
        std::copy(p->rct_pos.begin(), p->rct_pos.end(), h_pos);
        return AXSPMV_OK;
    }
    if (!adjoint_kind(p)) return fail(AXSPMV_E_UNSUPPORTED, "vbar_positions: plan %s has no value gradient", p->name);
    for (uint64_t k = 0; k < count; ++k) h_pos[k] = uint32_t(k);
    return AXSPMV_OK;
}
