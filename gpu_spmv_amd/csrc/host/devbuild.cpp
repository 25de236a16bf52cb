// Format builders on the device: parameter checks, sizing and allocation on
// the host; count, scan and fill kernels on the device (k_build.hip).  The
// arrays equal the host builders' (builders.cpp) bit for bit.
#include <cstdio>
#include <vector>

#include "../hip/launch.hpp"
#include "common.hpp"
#include "plan.hpp"

using namespace axspmv;

namespace axspmv {
// Scratch of one device build: the CSR check flag; for the counted layouts
// the per-row counts, their exclusive scan and rocPRIM's temporary storage.
struct BuildScratch {
    uint32_t *flag = nullptr;
    uint64_t *cnt = nullptr, *start = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
};
}  // namespace axspmv

namespace {

int check_dcsr(const axspmv_dcsr *a, const char *who) {
    if (!a) return fail(AXSPMV_E_INVALID, "%s: NULL matrix", who);
    if (!a->row || (a->nnz && (!a->col || !a->val))) return fail(AXSPMV_E_INVALID, "%s: NULL array", who);
    return AXSPMV_OK;
}

struct Events {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Events() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
    hipError_t create() {
        for (hipEvent_t &x : e) {
            const hipError_t r = hipEventCreate(&x);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
};

// Device memory a build allocates for itself, freed when the build ends.
struct Scratch {
    std::vector<void *> p;
    ~Scratch() {
        for (void *q : p) (void)hipFree(q);
    }
    void *alloc(size_t bytes) {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 8) != hipSuccess) return nullptr;
        p.push_back(q);
        return q;
    }
};

#define HIP_TRY(expr, what)                              \
    do {                                                 \
        const hipError_t e_ = (expr);                    \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// kind 0: AXC / AXT UNC th == 1 (count = tiles of hw), 1: AXT UNC th > 1
// (count = tile columns of th), 2 / 3: COM th == 1 / th > 1 (no count).
// w: the build's scratch (the check flag; for kinds 0 and 1 the counts, their
// scan and the scan's temporary storage).  own = true allocates the output
// arrays into f; own = false fills the caller's arrays in f, whose lengths
// must be the layout's.  Device memory is cleared with fill kernels.
int run_build(const axspmv_dcsr *a, const double *d_x, int kind, const launch::BuildGeo &g, bool axc,
              hipStream_t s, axspmv_dfmt *f, const BuildScratch &w, bool own) {
    const uint32_t n = a->nrows;
    const launch::DevCsr dc{n, a->ncols, a->nnz, a->row, a->col, a->val};
    const uint64_t ths = uint64_t(g.hw) * g.th;
    Events ev;
    HIP_TRY(ev.create(), "dev build: events");
    uint64_t total = 0;
    float ms_scan = 0.0f;
    // the CSR is checked on the device; its flag comes back with the scan
    // total, before anything reads col / val through the offsets
    uint32_t flag = 0;
    HIP_TRY(launch::fill_u32(w.flag, 1, 0u, s), "dev build: clear");
    HIP_TRY(launch::build_check(dc, w.flag, s), "dev build: check");
    if (kind <= 1) {
        size_t tmp_bytes = w.tmp_bytes;
        // counts of rows 0 .. n-1; the scan's n + 1st input is 0
        HIP_TRY(launch::fill_u32(reinterpret_cast<uint32_t *>(w.cnt + n), 2, 0u, s), "dev build: clear");
        HIP_TRY(hipEventRecord(ev.e[0], s), "dev build: event");
        HIP_TRY(launch::build_count(dc, kind == 0 ? g.hw : g.th, w.cnt, s), "dev build: count");
        HIP_TRY(launch::build_scan(w.cnt, w.start, size_t(n) + 1, w.tmp, tmp_bytes, s), "dev build: scan");
        HIP_TRY(hipEventRecord(ev.e[1], s), "dev build: event");
        HIP_TRY(hipMemcpyAsync(&total, w.start + n, 8, hipMemcpyDeviceToHost, s), "dev build: total");
    }
    HIP_TRY(hipMemcpyAsync(&flag, w.flag, 4, hipMemcpyDeviceToHost, s), "dev build: check flag");
    HIP_TRY(hipStreamSynchronize(s), "dev build: sync");
    if (flag)
        return fail(AXSPMV_E_INVALID, "dev build: invalid device CSR:%s%s%s%s", flag & 1 ? " row offsets decrease;" : "",
                    flag & 2 ? " row[0] != 0;" : "", flag & 4 ? " row[nrows] != nnz;" : "",
                    flag & 8 ? " a column >= ncols;" : "");
    if (kind <= 1) HIP_TRY(hipEventElapsedTime(&ms_scan, ev.e[0], ev.e[1]), "dev build: elapsed");
    uint64_t tileN, lenSEC, lenAX;
    if (kind == 0) {
        tileN = total;
        lenSEC = total;
        lenAX = total * 2 * g.hw;
    } else if (kind == 1) {
        tileN = ceil_div(total, g.hw);
        lenSEC = tileN * g.hw;
        lenAX = tileN * 2 * ths;
    } else {
        tileN = ceil_div(a->nnz, ths);
        lenSEC = tileN * ths;
        lenAX = tileN * 2 * ths;
    }
    if (lenAX >= (1ull << 32) || lenSEC >= (1ull << 32))
        return fail(AXSPMV_E_OVERFLOW, "dev build: lenAX %llu exceeds 32 bits", (unsigned long long)lenAX);
    if (axc) lenSEC = 0;
    const uint64_t nplane = axc ? lenAX / 2 + 8 : lenAX / 2;
    if (!own) {
        if (f->lenAX != lenAX || f->lenSEC != lenSEC || f->lenCON != n || f->lenMAPX != nplane)
            return fail(AXSPMV_E_INVALID,
                        "dev build: the arrays given hold lenAX %u, lenSEC %u, lenCON %u, lenMAPX %u; the layout needs "
                        "%llu, %llu, %u, %llu",
                        f->lenAX, f->lenSEC, f->lenCON, f->lenMAPX, (unsigned long long)lenAX,
                        (unsigned long long)lenSEC, n, (unsigned long long)nplane);
        if ((lenAX && (!f->ax || !f->plane)) || !f->con || (lenSEC && !f->sec))
            return fail(AXSPMV_E_INVALID, "dev build: NULL output array");
    }
    f->nrows = n;
    f->nnz = a->nnz;
    f->tileN = uint32_t(tileN);
    f->lenAX = uint32_t(lenAX);
    f->lenSEC = uint32_t(lenSEC);
    f->lenCON = n;
    f->lenMAPX = uint32_t(nplane);
    if (own && (hipMalloc(&f->ax, 8 * (lenAX ? lenAX : 1)) != hipSuccess ||
                hipMalloc(&f->plane, 4 * (nplane ? nplane : 1)) != hipSuccess ||
                hipMalloc(&f->con, 4ull * (n + 1)) != hipSuccess ||
                (!axc && hipMalloc(&f->sec, 4 * (lenSEC ? lenSEC : 1)) != hipSuccess)))
        return fail(AXSPMV_E_NOMEM, "dev build: device allocation failed");
    // zero fill: padding slots, headers and con[nrows] of the th > 1 layouts
    HIP_TRY(launch::fill_f64(f->ax, lenAX, 0.0, s), "dev build: clear");
    HIP_TRY(launch::fill_u32(f->plane, nplane, 0u, s), "dev build: clear");
    HIP_TRY(launch::fill_u32(f->con, size_t(n) + 1, 0u, s), "dev build: clear");
    if (!axc) HIP_TRY(launch::fill_u32(f->sec, lenSEC, 0u, s), "dev build: clear");
    HIP_TRY(hipEventRecord(ev.e[2], s), "dev build: event");
    const launch::BuildOut o{f->ax, f->plane, axc ? nullptr : f->sec, f->con};
    HIP_TRY(launch::build_fill(dc, kind, w.start, g, d_x, o, s), "dev build: fill");
    HIP_TRY(hipEventRecord(ev.e[3], s), "dev build: event");
    HIP_TRY(hipStreamSynchronize(s), "dev build: sync");
    float ms_fill = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms_fill, ev.e[2], ev.e[3]), "dev build: elapsed");
    f->ms = double(ms_scan) + double(ms_fill);
    return AXSPMV_OK;
}

// Temporary storage of rocPRIM's scan over n + 1 counts.
int scan_tmp_bytes(uint32_t n, hipStream_t s, size_t *bytes) {
    *bytes = 0;
    HIP_TRY(launch::build_scan(nullptr, nullptr, size_t(n) + 1, nullptr, *bytes, s), "dev build: scan size");
    return AXSPMV_OK;
}

// Parameters of an AXT device build: kind, header shift and name (the checks
// of axspmv_build_axt).
int axt_params(const axspmv_dcsr *a, uint32_t bs, uint32_t thw, uint32_t th, uint32_t mode, int *kind, uint32_t *log,
               char *name, size_t name_len) {
    if (thw == 0 || th == 0 || th > 1023) return fail(AXSPMV_E_INVALID, "dev_build_axt: need thw > 0 and 0 < th < 1024");
    if (mode != AXSPMV_AXT_UNC && mode != AXSPMV_AXT_COM)
        return fail(AXSPMV_E_INVALID, "dev_build_axt: mode must be UNC or COM");
    const bool com = mode == AXSPMV_AXT_COM;
    *log = 0;
    if (com && th == 1) {
        if (!is_pow2(bs) || bs < 2 || bs > 1024)
            return fail(AXSPMV_E_INVALID, "dev_build_axt: COM th==1 needs bs a power of two in [2,1024], got %u", bs);
        *log = ilog2(bs);
    } else if (com) {
        if (!is_pow2(th)) return fail(AXSPMV_E_INVALID, "dev_build_axt: COM needs th a power of two, got %u", th);
        *log = ilog2(th);
    }
    if (com && a->nrows && (uint64_t(a->nrows - 1) << *log) >= (1ull << 32))
        return fail(AXSPMV_E_OVERFLOW, "dev_build_axt: row << %u overflows the 32-bit header", *log);
    *kind = com ? (th == 1 ? 2 : 3) : (th == 1 ? 0 : 1);
    if (!com && th == 1)
        std::snprintf(name, name_len, "faxtuh1hw%u", thw);
    else if (!com)
        std::snprintf(name, name_len, "faxtuh%uhw%u", th, thw);
    else if (th == 1)
        std::snprintf(name, name_len, "faxtch1hw%ubs%u", thw, bs);
    else
        std::snprintf(name, name_len, "faxtch%uhw%u", th, thw);
    return AXSPMV_OK;
}

void set_axt_header(axspmv_dfmt *f, const char *name, uint32_t bs, uint32_t thw, uint32_t th, uint32_t mode,
                    uint32_t log) {
    std::snprintf(f->name, sizeof f->name, "%s", name);
    f->format = AXSPMV_FMT_AXT;
    f->mode = mode;
    f->hw = thw;
    f->th = th;
    f->bs = bs;
    f->log = log;
}

}  // namespace

extern "C" void axspmv_dfmt_free(axspmv_dfmt *f) {
    if (!f) return;
    if (f->ax) (void)hipFree(f->ax);
    if (f->sec) (void)hipFree(f->sec);
    if (f->con) (void)hipFree(f->con);
    if (f->plane) (void)hipFree(f->plane);
    std::free(f);
}

extern "C" int axspmv_dev_build_axt(const axspmv_dcsr *a, const double *d_x, uint32_t bs, uint32_t thw, uint32_t th,
                                    uint32_t mode, void *stream, axspmv_dfmt **out) {
    AX_TRY(check_dcsr(a, "dev_build_axt"));
    if (!out) return fail(AXSPMV_E_INVALID, "dev_build_axt: out is NULL");
    *out = nullptr;
    int kind = 0;
    uint32_t log = 0;
    char name[48];
    AX_TRY(axt_params(a, bs, thw, th, mode, &kind, &log, name, sizeof name));
    hipStream_t s = static_cast<hipStream_t>(stream);
    Scratch scratch;
    BuildScratch w;
    w.flag = static_cast<uint32_t *>(scratch.alloc(4));
    if (kind <= 1) {
        AX_TRY(scan_tmp_bytes(a->nrows, s, &w.tmp_bytes));
        w.cnt = static_cast<uint64_t *>(scratch.alloc(8ull * (a->nrows + 1)));
        w.start = static_cast<uint64_t *>(scratch.alloc(8ull * (a->nrows + 1)));
        w.tmp = scratch.alloc(w.tmp_bytes);
        if (!w.cnt || !w.start || !w.tmp) return fail(AXSPMV_E_NOMEM, "dev build: scratch allocation failed");
    }
    if (!w.flag) return fail(AXSPMV_E_NOMEM, "dev build: scratch allocation failed");
    auto *f = host_calloc<axspmv_dfmt>(1);
    if (!f) return AXSPMV_E_NOMEM;
    set_axt_header(f, name, bs, thw, th, mode, log);
    const int rc = run_build(a, d_x, kind, launch::BuildGeo{thw, th, bs, log}, false, s, f, w, true);
    if (rc) {
        axspmv_dfmt_free(f);
        return rc;
    }
    *out = f;
    return AXSPMV_OK;
}

extern "C" int axspmv_dev_build_axc(const axspmv_dcsr *a, const double *d_x, uint32_t hbs, void *stream,
                                    axspmv_dfmt **out) {
    AX_TRY(check_dcsr(a, "dev_build_axc"));
    if (!out || hbs == 0) return fail(AXSPMV_E_INVALID, "dev_build_axc: hbs must be > 0");
    *out = nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Scratch scratch;
    BuildScratch w;
    AX_TRY(scan_tmp_bytes(a->nrows, s, &w.tmp_bytes));
    w.flag = static_cast<uint32_t *>(scratch.alloc(4));
    w.cnt = static_cast<uint64_t *>(scratch.alloc(8ull * (a->nrows + 1)));
    w.start = static_cast<uint64_t *>(scratch.alloc(8ull * (a->nrows + 1)));
    w.tmp = scratch.alloc(w.tmp_bytes);
    if (!w.flag || !w.cnt || !w.start || !w.tmp) return fail(AXSPMV_E_NOMEM, "dev build: scratch allocation failed");
    auto *f = host_calloc<axspmv_dfmt>(1);
    if (!f) return AXSPMV_E_NOMEM;
    f->format = AXSPMV_FMT_AXC;
    f->hw = hbs;
    f->th = 1;
    std::snprintf(f->name, sizeof f->name, "faxc");
    const int rc = run_build(a, d_x, 0, launch::BuildGeo{hbs, 1, 0, 0}, true, s, f, w, true);
    if (rc) {
        axspmv_dfmt_free(f);
        return rc;
    }
    *out = f;
    return AXSPMV_OK;
}

namespace axspmv {

int dev_build_axt_into(const axspmv_dcsr *a, const double *d_x, uint32_t bs, uint32_t thw, uint32_t th, uint32_t mode,
                       hipStream_t s, char *scratch, size_t cap, axspmv_dfmt *f) {
    AX_TRY(check_dcsr(a, "dev_build_axt_noalloc"));
    if (!f) return fail(AXSPMV_E_INVALID, "dev_build_axt_noalloc: NULL descriptor");
    int kind = 0;
    uint32_t log = 0;
    char name[48];
    AX_TRY(axt_params(a, bs, thw, th, mode, &kind, &log, name, sizeof name));
    // flag | counts | starts | scan storage, each at a 256-byte boundary
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    BuildScratch w;
    size_t need = 256;
    if (kind <= 1) {
        AX_TRY(scan_tmp_bytes(a->nrows, s, &w.tmp_bytes));
        const size_t cnt_bytes = up(8 * (size_t(a->nrows) + 1));
        w.cnt = reinterpret_cast<uint64_t *>(scratch + need);
        w.start = reinterpret_cast<uint64_t *>(scratch + need + cnt_bytes);
        w.tmp = scratch + need + 2 * cnt_bytes;
        need += 2 * cnt_bytes + up(w.tmp_bytes);
    }
    if (need > cap)
        return fail(AXSPMV_E_UNSUPPORTED,
                    "dev_build_axt_noalloc: %u rows need %zu bytes of scratch, the static workspace holds %zu "
                    "(axspmv_dev_build_axt has no such limit)",
                    a->nrows, need, cap);
    w.flag = reinterpret_cast<uint32_t *>(scratch);
    // *f changes only when the build succeeds
    axspmv_dfmt g = *f;
    set_axt_header(&g, name, bs, thw, th, mode, log);
    const int rc = run_build(a, d_x, kind, launch::BuildGeo{thw, th, bs, log}, false, s, &g, w, false);
    if (rc == AXSPMV_OK) *f = g;
    return rc;
}

}  // namespace axspmv
