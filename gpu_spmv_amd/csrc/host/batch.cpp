// Batched plans (many matrices, one persistent launch) and the autotuner.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>

#include "plan.hpp"

using namespace axspmv;

struct axspmv_batch {
    std::vector<void *> bufs;
    launch::BatchArgs args{};
    axspmv_batch_info info{};
    axspmv_plan *rct = nullptr;  // RCT batches: the plan of the block-diagonal matrix
    void *dalloc(size_t bytes) {
        void *p = nullptr;
        if (bytes == 0) bytes = 256;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            fail(AXSPMV_E_NOMEM, "hipMalloc of %zu bytes failed", bytes);
            return nullptr;
        }
        bufs.push_back(p);
        info.device_bytes += bytes;
        return p;
    }
    ~axspmv_batch() {
        for (void *p : bufs) (void)hipFree(p);
        if (rct) axspmv_plan_destroy(rct);
    }
    // the batch's SpMV on stream st
    hipError_t launch_on(const double *x, double *y, hipStream_t st) const {
        return rct ? plan_launch(rct, x, y, st) : launch::batch(args, x, y, st);
    }
};

namespace {

// Waves of the persistent grid: 16 per CU.
uint32_t persistent_waves() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    return uint32_t(cus) * 16;
}

int h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return AXSPMV_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "batch upload");
}

// check_csr_bounds of batch member i, its message prefixed with the index.
int check_member(const axspmv_csr *a, uint32_t i) {
    char who[64];
    std::snprintf(who, sizeof who, "batch_create: matrix %u", i);
    return check_csr_bounds(a, who);
}

// Algorithmic bytes of one SpMV of a: values and column indices, row
// pointer, x and y once.
uint64_t alg_bytes(const axspmv_csr *a, size_t vs) {
    return (vs + 4) * uint64_t(a->nnz) + 4ull * (a->nrows + 1) + 8ull * a->ncols + 8ull * a->nrows;
}

// RCT batch: the block-diagonal matrix diag(A_0, ..., A_{count-1}) -- x and
// y concatenated in batch order, as for every batch -- as one RCT plan.  Its
// row blocks balance nonzeros across matrix boundaries, and banded members
// keep it at one column slice with packed keys.
int batch_create_rct(const axspmv_csr *const *mats, uint32_t count, const axspmv_opts *opts, int per_matrix,
                     void *stream, axspmv_batch **out) {
    for (uint32_t i = 1; per_matrix && i < count; ++i)
        if (std::memcmp(&opts[i], &opts[0], sizeof(axspmv_opts)) != 0)
            return fail(AXSPMV_E_UNSUPPORTED, "batch_create: an RCT batch takes one option set (matrix %u differs)", i);
    uint64_t nrows = 0, ncols = 0, nnz = 0, alg = 0;
    const size_t vs = opts[0].vtype == AXSPMV_F32 ? 4 : 8;
    for (uint32_t i = 0; i < count; ++i) {
        const axspmv_csr *a = mats[i];
        AX_TRY(check_member(a, i));
        nrows += a->nrows;
        ncols += a->ncols;
        nnz += a->nnz;
        alg += alg_bytes(a, vs);
    }
    if (nrows >= (1ull << 32) || ncols >= (1ull << 32) || nnz >= (1ull << 32))
        return fail(AXSPMV_E_OVERFLOW, "batch_create: the block-diagonal matrix exceeds 32-bit sizes");
    std::vector<uint32_t> row(nrows + 1), col(nnz ? nnz : 1);
    std::vector<double> val(nnz ? nnz : 1);
    uint64_t r0 = 0, c0 = 0, e0 = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const axspmv_csr *a = mats[i];
        for (uint32_t r = 0; r < a->nrows; ++r) row[r0 + r] = uint32_t(e0 + a->row[r]);
        for (uint32_t j = 0; j < a->nnz; ++j) col[e0 + j] = uint32_t(c0 + a->col[j]);
        if (a->nnz) std::memcpy(val.data() + e0, a->val, 8ull * a->nnz);
        r0 += a->nrows;
        c0 += a->ncols;
        e0 += a->nnz;
    }
    row[nrows] = uint32_t(nnz);
    axspmv_csr bd{};
    bd.nrows = uint32_t(nrows);
    bd.ncols = uint32_t(ncols);
    bd.nnz = uint32_t(nnz);
    bd.row = row.data();
    bd.col = col.data();
    bd.val = val.data();
    axspmv_opts o = opts[0];
    o.flags &= ~AXSPMV_PLAN_TRANSPOSE;
    auto b = std::make_unique<axspmv_batch>();
    AX_TRY(axspmv_plan_create(&bd, &o, stream, &b->rct));
    axspmv_plan_info pi;
    AX_TRY(axspmv_plan_get_info(b->rct, &pi));
    b->info.nrows = nrows;
    b->info.ncols = ncols;
    b->info.nnz = nnz;
    b->info.slots = pi.slots;
    b->info.alg_bytes = alg;
    b->info.device_bytes = pi.device_bytes;
    b->info.count = count;
    b->info.nitems = pi.nblocks;  // tiles
    b->info.nwaves = 0;
    *out = b.release();
    return AXSPMV_OK;
}

}  // namespace

extern "C" int axspmv_batch_create(const axspmv_csr *const *mats, uint32_t count, const axspmv_opts *opts,
                                   int per_matrix, void *stream, axspmv_batch **out) {
    if (!mats || !opts || !out || count == 0) return fail(AXSPMV_E_INVALID, "batch_create: bad argument");
    *out = nullptr;
    for (uint32_t i = 0; i < (per_matrix ? count : 1u); ++i)
        if (opts[i].flags & AXSPMV_PLAN_REORDER)
            return fail(AXSPMV_E_UNSUPPORTED, "batch_create: AXSPMV_PLAN_REORDER is for single plans");
    if (opts[0].format == AXSPMV_FMT_RCT) return batch_create_rct(mats, count, opts, per_matrix, stream, out);
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto b = std::make_unique<axspmv_batch>();
    const uint32_t vtype = opts[0].vtype;
    const size_t vs = vtype == AXSPMV_F32 ? 4 : 8;
    std::vector<launch::BatchMat> desc(count);
    std::vector<launch::BatchItem> items;
    std::vector<launch::Window> wins;  // column window of every item
    std::vector<uint32_t> empty;
    std::vector<double> vcat;
    std::vector<float> vcat32;
    std::vector<uint32_t> ccat, scat;
    uint64_t xoff = 0, yoff = 0;
    // A matrix listed more than once (the tuner repeats a group's members)
    // is built once per tile shape; builds of the others are freed as soon
    // as they are concatenated.
    using AxtPtr = std::unique_ptr<axspmv_axt, void (*)(axspmv_axt *)>;
    std::map<const axspmv_csr *, uint32_t> seen;
    for (uint32_t i = 0; i < count; ++i) seen[mats[i]]++;
    std::map<std::tuple<const axspmv_csr *, uint32_t, uint32_t>, AxtPtr> built;
    for (uint32_t i = 0; i < count; ++i) {
        const axspmv_csr *a = mats[i];
        axspmv_opts o = opts[per_matrix ? i : 0];
        AX_TRY(check_member(a, i));
        if (o.format != AXSPMV_FMT_AXT || o.mode != AXSPMV_AXT_UNC)
            return fail(AXSPMV_E_UNSUPPORTED, "batch_create: matrix %u: batches use the AXT UNC layout", i);
        if (o.vtype != vtype) return fail(AXSPMV_E_INVALID, "batch_create: one value type per batch");
        if (o.slices > 1)
            return fail(AXSPMV_E_UNSUPPORTED, "batch_create: matrix %u: AXT column slices are for single plans", i);
        if (o.th == 0) o.th = 1;
        if (o.thw == 0) o.thw = o.th == 1 ? 4 : 64;
        if (!is_pow2(o.thw) || o.thw > 64) return fail(AXSPMV_E_UNSUPPORTED, "batch_create: thw must be a power of two <= 64");
        const auto key = std::make_tuple(a, o.thw, o.th);
        auto hit = built.find(key);
        AxtPtr own(nullptr, axspmv_axt_free);
        const axspmv_axt *m = nullptr;
        if (hit != built.end()) {
            m = hit->second.get();
        } else {
            axspmv_axt *fresh = nullptr;
            AX_TRY(axspmv_build_axt(a, nullptr, 256, o.thw, o.th, AXSPMV_AXT_UNC,
                                    AXSPMV_BUILD_PLANES | AXSPMV_BUILD_PAD_COL, &fresh, nullptr));
            own.reset(fresh);
            m = fresh;
            if (seen[a] > 1) built.emplace(key, std::move(own));
        }
        // keep every matrix's planes 64-slot aligned (vector loads)
        const uint64_t voff = vtype == AXSPMV_F32 ? vcat32.size() : vcat.size();
        const uint64_t nslots = m->lenAX / 2;
        if (vtype == AXSPMV_F32)
            vcat32.insert(vcat32.end(), m->vplane, m->vplane + nslots);
        else
            vcat.insert(vcat.end(), m->vplane, m->vplane + nslots);
        ccat.insert(ccat.end(), m->cplane, m->cplane + nslots);
        const uint64_t pad = (64 - nslots % 64) % 64;
        if (vtype == AXSPMV_F32)
            vcat32.insert(vcat32.end(), pad, 0.0f);
        else
            vcat.insert(vcat.end(), pad, 0.0);
        ccat.insert(ccat.end(), pad, 0u);
        const uint64_t soff = scat.size();
        scat.insert(scat.end(), m->sec, m->sec + m->lenSEC);
        launch::BatchMat d{};
        d.voff = voff;
        d.soff = soff;
        if (xoff + a->ncols >= (1ull << 32) || yoff + a->nrows >= (1ull << 32))
            return fail(AXSPMV_E_OVERFLOW, "batch_create: concatenated vectors exceed 32-bit offsets");
        d.xoff = uint32_t(xoff);
        d.yoff = uint32_t(yoff);
        const uint32_t K = o.th == 1 ? std::min<uint32_t>(4, o.thw) : 1;
        d.kind = o.th == 1 ? (K == 4 ? 0u : K == 2 ? 1u : 2u) : 3u;
        d.ushift = o.th == 1 ? ilog2(o.thw / K) : 0;
        d.thw_log = ilog2(o.thw);
        d.th = o.th;
        desc[i] = d;
        std::vector<uint32_t> units(a->nrows);
        for (uint32_t r = 0; r < a->nrows; ++r) {
            const uint32_t rl = a->row[r + 1] - a->row[r];
            if (rl == 0) empty.push_back(uint32_t(yoff) + r);
            units[r] = o.th == 1 ? uint32_t(ceil_div(rl, o.thw) * (o.thw / K)) : uint32_t(ceil_div(rl, o.th));
        }
        const std::vector<launch::Block> blocks = row_blocks(units, o.budget ? o.budget : 256);
        const std::vector<launch::Window> w = block_windows(a, units, blocks);
        for (size_t k = 0; k < blocks.size(); ++k) {
            items.push_back({i, blocks[k].beg, blocks[k].end, 0});
            wins.push_back(w[k]);
        }
        b->info.nrows += a->nrows;
        b->info.ncols += a->ncols;
        b->info.nnz += a->nnz;
        b->info.slots += nslots;
        b->info.alg_bytes += alg_bytes(a, vs);
        xoff += a->ncols;
        yoff += a->nrows;
    }
    // longest blocks first across the whole batch (windows follow their items)
    std::vector<uint32_t> order(items.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t p, uint32_t q) {
        return items[p].end - items[p].beg > items[q].end - items[q].beg;
    });
    {
        std::vector<launch::BatchItem> it2(items.size());
        std::vector<launch::Window> w2(wins.size());
        for (size_t k = 0; k < order.size(); ++k) {
            it2[k] = items[order[k]];
            w2[k] = wins[order[k]];
        }
        items.swap(it2);
        wins.swap(w2);
    }
    const bool stage = (opts[0].flags & AXSPMV_PLAN_STAGE_X) && worth_staging(wins);
    const uint64_t nsl = ccat.size();
    void *dv = b->dalloc(vs * nsl);
    auto *dc = static_cast<uint32_t *>(b->dalloc(4 * nsl));
    auto *ds = static_cast<uint32_t *>(b->dalloc(4 * scat.size()));
    auto *dm = static_cast<launch::BatchMat *>(b->dalloc(sizeof(launch::BatchMat) * desc.size()));
    auto *di = static_cast<launch::BatchItem *>(b->dalloc(sizeof(launch::BatchItem) * items.size()));
    auto *de = static_cast<uint32_t *>(b->dalloc(4 * empty.size()));
    auto *dw = stage ? static_cast<launch::Window *>(b->dalloc(sizeof(launch::Window) * wins.size())) : nullptr;
    if (!dv || !dc || !ds || !dm || !di || !de || (stage && !dw)) return AXSPMV_E_NOMEM;
    if (stage) AX_TRY(h2d(dw, wins.data(), sizeof(launch::Window) * wins.size(), s));
    AX_TRY(h2d(dv, vtype == AXSPMV_F32 ? static_cast<const void *>(vcat32.data()) : vcat.data(), vs * nsl, s));
    AX_TRY(h2d(dc, ccat.data(), 4 * nsl, s));
    AX_TRY(h2d(ds, scat.data(), 4 * scat.size(), s));
    AX_TRY(h2d(dm, desc.data(), sizeof(launch::BatchMat) * desc.size(), s));
    AX_TRY(h2d(di, items.data(), sizeof(launch::BatchItem) * items.size(), s));
    AX_TRY(h2d(de, empty.data(), 4 * empty.size(), s));
    launch::BatchArgs &g = b->args;
    g.vplane = dv;
    g.cplane = dc;
    g.sec = ds;
    g.mats = dm;
    g.items = di;
    g.nitems = uint32_t(items.size());
    g.nwaves = std::max<uint32_t>(1, std::min<uint32_t>(g.nitems, persistent_waves()));
    g.vtype = vtype;
    g.empty = de;
    g.nempty = uint32_t(empty.size());
    g.win = dw;
    b->info.count = count;
    b->info.nitems = g.nitems;
    b->info.nwaves = g.nwaves;
    *out = b.release();
    return AXSPMV_OK;
}

extern "C" int axspmv_batch_destroy(axspmv_batch *b) {
    delete b;
    return AXSPMV_OK;
}

extern "C" int axspmv_batch_get_info(const axspmv_batch *b, axspmv_batch_info *info) {
    if (!b || !info) return fail(AXSPMV_E_INVALID, "batch_get_info: NULL argument");
    *info = b->info;
    return AXSPMV_OK;
}

extern "C" int axspmv_batch_spmv(axspmv_batch *b, const double *d_x, double *d_y, void *stream) {
    if (!b || !d_x || !d_y) return fail(AXSPMV_E_INVALID, "batch_spmv: NULL argument");
    const hipError_t e = b->launch_on(d_x, d_y, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "batch launch");
}

extern "C" int axspmv_batch_time(axspmv_batch *b, const double *d_x, double *d_y, uint32_t iters, uint32_t warmup,
                                 uint32_t mode, uint32_t graph_batch, void *stream, axspmv_timing *out) {
    if (!b || !d_x || !d_y) return fail(AXSPMV_E_INVALID, "batch_time: NULL argument");
    return time_launches([b, d_x, d_y](hipStream_t st) { return b->launch_on(d_x, d_y, st); }, iters, warmup, mode,
                         graph_batch, stream, out);
}

// ---------------------------------------------------------------- tuner
namespace {

// Built-in candidates.  Batches take AXT UNC per shape group and RCT (the
// block-diagonal plan) for the whole batch; single plans also try AXT UNC
// th = 1 in 2, 4 and 8 column slices, K1 and RCT.  SpMM: AXT
// k-tiles, and RCT tiles with 64 KB of rows x k-tile sums (one column slice
// and the default slicing) or 32 KB (one slice).
std::vector<axspmv_opts> default_candidates(uint32_t op, bool for_batch = false) {
    std::vector<axspmv_opts> c;
    axspmv_opts o;
    axspmv_opts_default(AXSPMV_FMT_AXT, &o);
    if (op == AXSPMV_OP_SPMM) {
        for (uint32_t kt : {8u, 16u, 32u, 64u}) {
            o.ktile = kt;
            c.push_back(o);
        }
        axspmv_opts_default(AXSPMV_FMT_RCT, &o);
        for (uint32_t kt : {8u, 16u, 32u}) {
            o.ktile = kt;
            o.tile_rows = 8192 / kt;  // 64 KB of sums (ktile 8: also 2048 rows)
            for (uint32_t sl : {1u, 0u}) {
                o.slices = sl;
                c.push_back(o);
                if (kt == 8) {
                    axspmv_opts w = o;
                    w.tile_rows = 2048;
                    c.push_back(w);
                }
            }
            // half the height, one slice (k = 32 on the 2^20-row power-law
            // matrix: 256 rows at ktile 16 took 1019 us, 512 rows 1060-1080)
            axspmv_opts h = o;
            h.tile_rows = 4096 / kt;
            h.slices = 1;
            c.push_back(h);
        }
        return c;
    }
    for (uint32_t budget : {64u, 256u}) {
        o.budget = budget;
        for (uint32_t thw : {2u, 4u, 8u, 16u}) {
            o.th = 1;
            o.thw = thw;
            c.push_back(o);
        }
        for (uint32_t th : {2u, 4u, 8u}) {
            o.th = th;
            o.thw = 64;
            c.push_back(o);
        }
    }
    if (for_batch) {
        axspmv_opts_default(AXSPMV_FMT_RCT, &o);
        c.push_back(o);
        return c;
    }
    // the AXT UNC th = 1 layout per column slice (two slots per tile)
    o.budget = 64;
    o.th = 1;
    o.thw = 2;
    for (uint32_t sl : {2u, 4u, 8u}) {
        o.slices = sl;
        c.push_back(o);
    }
    axspmv_opts_default(AXSPMV_FMT_K1, &o);
    c.push_back(o);
    axspmv_opts_default(AXSPMV_FMT_RCT, &o);
    for (uint32_t block : {0u, 256u, 1024u}) {
        o.block = block;
        c.push_back(o);
    }
    // renumbered by reverse Cuthill-McKee (square matrices only; where the
    // renumbering brings nothing the plan is the plain one, timed again)
    o.block = 0;
    o.flags = AXSPMV_PLAN_REORDER;
    c.push_back(o);
    axspmv_opts_default(AXSPMV_FMT_AXT, &o);
    o.flags = AXSPMV_PLAN_REORDER;
    c.push_back(o);
    return c;
}

}  // namespace

extern "C" int axspmv_autotune_defaults(uint32_t op, axspmv_opts *out, uint32_t cap, uint32_t *count) {
    if (!count || (op != AXSPMV_OP_SPMV && op != AXSPMV_OP_SPMM) || (cap && !out))
        return fail(AXSPMV_E_INVALID, "autotune_defaults: bad argument");
    const std::vector<axspmv_opts> cs = default_candidates(op);
    *count = uint32_t(cs.size());
    for (uint32_t i = 0; i < cap && i < cs.size(); ++i) out[i] = cs[i];
    return AXSPMV_OK;
}

extern "C" int axspmv_autotune(const axspmv_csr *a, uint32_t op, uint32_t k, const axspmv_opts *cand, uint32_t ncand,
                               uint32_t iters, void *stream, uint32_t *best, double *times_ms) {
    if (!a || !best || (op != AXSPMV_OP_SPMV && op != AXSPMV_OP_SPMM) || (op == AXSPMV_OP_SPMM && k == 0))
        return fail(AXSPMV_E_INVALID, "autotune: bad argument");
    std::vector<axspmv_opts> cs = cand ? std::vector<axspmv_opts>(cand, cand + ncand) : default_candidates(op);
    if (cs.empty()) return fail(AXSPMV_E_INVALID, "autotune: no candidates");
    if (!cand && times_ms && ncand < cs.size())
        return fail(AXSPMV_E_INVALID, "autotune: times_ms holds %u entries, the built-in set has %zu", ncand,
                    cs.size());
    if (iters == 0) iters = 20;
    const uint64_t kk = op == AXSPMV_OP_SPMM ? k : 1;
    void *dx = nullptr, *dy = nullptr;
    if (hipMalloc(&dx, 8 * std::max<uint64_t>(1, uint64_t(a->ncols) * kk)) != hipSuccess ||
        hipMalloc(&dy, 8 * std::max<uint64_t>(1, uint64_t(a->nrows) * kk)) != hipSuccess) {
        if (dx) (void)hipFree(dx);
        return fail(AXSPMV_E_NOMEM, "autotune: vector allocation failed");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    // x = 1 (the timing does not depend on the values)
    if (launch::fill_f64(static_cast<double *>(dx), std::max<uint64_t>(1, uint64_t(a->ncols) * kk), 1.0, s) !=
        hipSuccess) {
        (void)hipFree(dx);
        (void)hipFree(dy);
        return fail(AXSPMV_E_HIP, "autotune: fill launch failed");
    }
    double best_t = INFINITY;
    uint32_t best_i = 0;
    for (uint32_t i = 0; i < cs.size(); ++i) {
        double t = -1.0;
        axspmv_plan *p = nullptr;
        if (axspmv_plan_create(a, &cs[i], stream, &p) == AXSPMV_OK) {
            axspmv_operands o{};
            o.x = o.X = static_cast<const double *>(dx);
            o.y = o.Y = static_cast<double *>(dy);
            o.k = k;
            axspmv_timing tm;
            if (axspmv_time_op(p, op, &o, iters, 2, AXSPMV_TIME_REGION, 1, stream, &tm) == AXSPMV_OK) t = tm.mean_ms;
            axspmv_plan_destroy(p);
        }
        if (times_ms) times_ms[i] = t;
        if (t >= 0 && t < best_t) {
            best_t = t;
            best_i = i;
        }
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    clear_error();
    if (!(best_t < INFINITY)) return fail(AXSPMV_E_UNSUPPORTED, "autotune: no candidate ran");
    *best = best_i;
    return AXSPMV_OK;
}

namespace {

// Nonzeros of the batch a group is tuned on: enough to fill the persistent
// grid several times over, so each candidate is timed at the throughput its
// matrices see inside a large batch, not at a lone small launch's latency.
constexpr uint64_t TUNE_BATCH_NNZ = 4ull << 20;
constexpr uint32_t TUNE_BATCH_MAX = 4096;

// Index of the fastest candidate on a batch of the group's members (repeated
// until it holds TUNE_BATCH_NNZ nonzeros).
int tune_group(const axspmv_csr *const *mats, const std::vector<uint32_t> &members,
               const std::vector<axspmv_opts> &cs, uint32_t iters, void *stream, uint32_t *best) {
    std::vector<const axspmv_csr *> sub;
    uint64_t nnz = 0, nrows = 0, ncols = 0;
    for (size_t i = 0; sub.size() < TUNE_BATCH_MAX && nnz < TUNE_BATCH_NNZ; ++i) {
        if (i == members.size() && nnz == 0) break;  // empty matrices only: once is enough
        const axspmv_csr *a = mats[members[i % members.size()]];
        sub.push_back(a);
        nnz += a->nnz;
        nrows += a->nrows;
        ncols += a->ncols;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    void *dx = nullptr, *dy = nullptr;
    if (hipMalloc(&dx, 8 * std::max<uint64_t>(1, ncols)) != hipSuccess ||
        hipMalloc(&dy, 8 * std::max<uint64_t>(1, nrows)) != hipSuccess) {
        if (dx) (void)hipFree(dx);
        return fail(AXSPMV_E_NOMEM, "autotune_batch: vector allocation failed");
    }
    int rc = launch::fill_f64(static_cast<double *>(dx), std::max<uint64_t>(1, ncols), 1.0, s) == hipSuccess
                 ? AXSPMV_OK
                 : fail(AXSPMV_E_HIP, "autotune_batch: fill launch failed");
    double best_t = INFINITY;
    *best = 0;
    for (uint32_t c = 0; rc == AXSPMV_OK && c < cs.size(); ++c) {
        axspmv_batch *b = nullptr;
        if (axspmv_batch_create(sub.data(), uint32_t(sub.size()), &cs[c], 0, stream, &b) != AXSPMV_OK) continue;
        axspmv_timing tm;
        if (axspmv_batch_time(b, static_cast<double *>(dx), static_cast<double *>(dy), iters, 2, AXSPMV_TIME_REGION,
                              1, stream, &tm) == AXSPMV_OK &&
            tm.mean_ms < best_t) {
            best_t = tm.mean_ms;
            *best = c;
        }
        axspmv_batch_destroy(b);
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    if (rc == AXSPMV_OK && !std::isfinite(best_t)) rc = fail(AXSPMV_E_UNSUPPORTED, "autotune_batch: no candidate ran");
    return rc;
}

}  // namespace

extern "C" int axspmv_autotune_batch(const axspmv_csr *const *mats, uint32_t count, const axspmv_opts *cand,
                                     uint32_t ncand, uint32_t iters, void *stream, axspmv_opts *out_opts,
                                     uint32_t *ngroups) {
    if (!mats || !out_opts || count == 0) return fail(AXSPMV_E_INVALID, "autotune_batch: bad argument");
    std::vector<axspmv_opts> cs =
        cand ? std::vector<axspmv_opts>(cand, cand + ncand) : default_candidates(AXSPMV_OP_SPMV, true);
    if (cs.empty()) return fail(AXSPMV_E_INVALID, "autotune_batch: no candidates");
    if (iters == 0) iters = 20;
    // AXT candidates are chosen per shape group, RCT ones for the whole batch
    std::vector<axspmv_opts> axt, rct;
    for (const axspmv_opts &o : cs) {
        if (o.format == AXSPMV_FMT_AXT)
            axt.push_back(o);
        else if (o.format == AXSPMV_FMT_RCT)
            rct.push_back(o);
        else
            return fail(AXSPMV_E_UNSUPPORTED, "autotune_batch: candidates must be AXT or RCT, got format %u", o.format);
    }
    std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> groups;  // shape -> members
    uint64_t nrows = 0, ncols = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const axspmv_csr *a = mats[i];
        if (!a) return fail(AXSPMV_E_INVALID, "autotune_batch: matrix %u is NULL", i);
        const uint32_t lg = a->nrows ? ilog2(a->nrows) : 0;
        const double mean = a->nrows ? double(a->nnz) / a->nrows : 0.0;
        groups[std::make_pair(lg, uint32_t(mean / 4.0))].push_back(i);
        nrows += a->nrows;
        ncols += a->ncols;
    }
    for (const auto &g : groups) {
        uint32_t best = 0;
        if (!axt.empty()) AX_TRY(tune_group(mats, g.second, axt, iters, stream, &best));
        for (uint32_t i : g.second) out_opts[i] = axt.empty() ? rct[0] : axt[best];
    }
    uint32_t chosen_sets = uint32_t(groups.size());
    if (!rct.empty()) {
        // The whole batch, timed with the per-group AXT choice and with each
        // RCT option set (the block-diagonal plan); the fastest wins.
        hipStream_t s = static_cast<hipStream_t>(stream);
        void *dx = nullptr, *dy = nullptr;
        if (hipMalloc(&dx, 8 * std::max<uint64_t>(1, ncols)) != hipSuccess ||
            hipMalloc(&dy, 8 * std::max<uint64_t>(1, nrows)) != hipSuccess) {
            if (dx) (void)hipFree(dx);
            return fail(AXSPMV_E_NOMEM, "autotune_batch: vector allocation failed");
        }
        struct Free {
            void *a, *b;
            ~Free() {
                (void)hipFree(a);
                (void)hipFree(b);
            }
        } guard{dx, dy};
        if (launch::fill_f64(static_cast<double *>(dx), std::max<uint64_t>(1, ncols), 1.0, s) != hipSuccess)
            return fail(AXSPMV_E_HIP, "autotune_batch: fill launch failed");
        auto time_batch = [&](const axspmv_opts *o, int per_matrix, double *ms) {
            axspmv_batch *b = nullptr;
            AX_TRY(axspmv_batch_create(mats, count, o, per_matrix, stream, &b));
            axspmv_timing tm;
            const int rc = axspmv_batch_time(b, static_cast<double *>(dx), static_cast<double *>(dy), iters, 2,
                                             AXSPMV_TIME_REGION, 1, stream, &tm);
            axspmv_batch_destroy(b);
            if (rc == AXSPMV_OK) *ms = tm.mean_ms;
            return rc;
        };
        double best = INFINITY;
        int which = -1;  // -1: the per-group AXT choice
        if (!axt.empty()) AX_TRY(time_batch(out_opts, 1, &best));
        for (size_t r = 0; r < rct.size(); ++r) {
            double ms = INFINITY;
            if (time_batch(&rct[r], 0, &ms) == AXSPMV_OK && ms < best) {
                best = ms;
                which = int(r);
            }
        }
        if (!std::isfinite(best)) return fail(AXSPMV_E_UNSUPPORTED, "autotune_batch: no candidate ran");
        if (which >= 0) {
            for (uint32_t i = 0; i < count; ++i) out_opts[i] = rct[which];
            chosen_sets = 1;
        }
    }
    if (ngroups) *ngroups = chosen_sets;
    return AXSPMV_OK;
}

extern "C" int axspmv_batch_graph(axspmv_batch *b, const double *d_x, double *d_y, uint32_t repeats,
                                  axspmv_graph **out) {
    if (!b || !d_x || !d_y) return fail(AXSPMV_E_INVALID, "batch_graph: NULL argument");
    return capture_graph([b, d_x, d_y](hipStream_t st) { return b->launch_on(d_x, d_y, st); }, repeats, out);
}
