// Format builders: K1, AXC and AXT (UNC/COM x th==1/th>1).
//
// The arrays are the reference's layouts (cudaSpmv.cu:708-846, 933-1062,
// 1154-1566) at whatever lane parameters the caller picks (32-lane for
// comparison with the reference, 64-lane for the wave64 kernels).  Instead
// of the reference's per-row walking pointers, every element's slot is
// computed in closed form, so rows are filled independently in parallel:
//
//   AXC / AXT UNC th==1   element k of row r -> slot S_r + k       (plane)
//                         where S_r = (first slot of row r) = con[r]/2
//   AXT UNC th>1          global column g = c_r + k/th, level f = k%th
//                         plane slot (g/thw)*thw*th + f*thw + g%thw
//   AXT COM th==1         plane slot = CSR position p
//   AXT COM th>1          p -> tile p/(thw*th), level p%th, column
//                         (p%(thw*th))/th; slot tile*thw*th + level*thw + column
//
// The interleaved ax[] slot of plane slot q is (q/thw)*2*thw + q%thw (value)
// and that plus thw (gathered x value).
#include <algorithm>
#include <numeric>
#include <string>

#include "common.hpp"

using namespace axspmv;

namespace {

inline uint64_t ax_of_plane(uint64_t q, uint32_t hw) { return (q / hw) * 2 * hw + q % hw; }

// The builders index x and their tables by column: offsets and columns are
// checked first.
int check_input(const axspmv_csr *a, const char *who) { return check_csr_bounds(a, who); }

void set_fd(axspmv_fmtdata *fd, const char *name, double mfp, double beta, double ct) {
    if (!fd) return;
    std::snprintf(fd->name, sizeof fd->name, "%s", name);
    fd->mfp = mfp;
    fd->beta = beta;
    fd->ct = ct;
}

// Row lengths from the row pointer (the reference reads rl[] filled by
// printMatrixStats; computing it here removes that hidden dependency).
inline uint32_t rlen(const axspmv_csr *a, uint32_t r) { return a->row[r + 1] - a->row[r]; }

}  // namespace

// ------------------------------------------------------------------- K1
extern "C" int axspmv_build_k1(const axspmv_csr *a, uint32_t chunk, int reorder, axspmv_k1 **out, axspmv_fmtdata *fd) {
    AX_TRY(check_input(a, "build_k1"));
    if (!out || chunk == 0) return fail(AXSPMV_E_INVALID, "build_k1: chunk must be > 0");
    *out = nullptr;
    const double t0 = now_s();
    const uint32_t n = a->nrows;
    const uint32_t nchunk = uint32_t(ceil_div(n, chunk));
    auto *m = host_calloc<axspmv_k1>(1);
    if (!m) return AXSPMV_E_NOMEM;
    m->nrows = n;
    m->nnz = a->nnz;
    m->chunk = chunk;
    m->chunkNum = nchunk;
    m->reorder = reorder ? 1 : 0;
    m->permi = host_calloc<uint32_t>((size_t(nchunk) + 1) * chunk);
    m->nmc = host_calloc<uint32_t>(nchunk);
    m->chp = host_calloc<uint32_t>(nchunk);
    if (!m->permi || !m->nmc || !m->chp) {
        axspmv_k1_free(m);
        return AXSPMV_E_NOMEM;
    }
    std::iota(m->permi, m->permi + n, 0u);
    if (reorder)  // length descending, ties by ascending row index
        std::stable_sort(m->permi, m->permi + n, [a](uint32_t x, uint32_t y) { return rlen(a, x) > rlen(a, y); });
    uint64_t len = 0;
    for (uint32_t c = 0; c < nchunk; ++c) {
        uint32_t mx = 0;
        const uint32_t r_end = std::min<uint64_t>(uint64_t(c + 1) * chunk, n);
        if (reorder)
            mx = rlen(a, m->permi[uint64_t(c) * chunk]);
        else
            for (uint32_t i = c * chunk; i < r_end; ++i) mx = std::max(mx, rlen(a, m->permi[i]));
        m->nmc[c] = mx;
        if (len >= (1ull << 32)) break;
        m->chp[c] = uint32_t(len);
        len += uint64_t(chunk) * mx;
    }
    if (len >= (1ull << 32)) {
        axspmv_k1_free(m);
        return fail(AXSPMV_E_OVERFLOW, "build_k1: lenVC exceeds 32 bits");
    }
    m->lenVC = uint32_t(len);
    m->val = host_calloc<double>(len);
    m->col = host_calloc<uint32_t>(len);
    if (!m->val || !m->col) {
        axspmv_k1_free(m);
        return AXSPMV_E_NOMEM;
    }
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = i / chunk, lane = i % chunk, r = m->permi[i];
        const uint32_t b = a->row[r], l = rlen(a, r);
        for (uint32_t k = 0; k < l; ++k) {
            const uint64_t q = uint64_t(m->chp[c]) + uint64_t(k) * chunk + lane;
            m->val[q] = a->val[b + k];
            m->col[q] = a->col[b + k];
        }
    }
    const double mfp = 8.0 * nchunk + 12.0 * len + 4.0 * n + 8.0 * n;
    set_fd(fd, reorder ? "fk1" : "fk1nat", mfp, len ? double(a->nnz) / double(len) : 0.0, now_s() - t0);
    *out = m;
    return AXSPMV_OK;
}

extern "C" void axspmv_k1_free(axspmv_k1 *m) {
    if (!m) return;
    std::free(m->permi);
    std::free(m->nmc);
    std::free(m->chp);
    std::free(m->val);
    std::free(m->col);
    std::free(m);
}

// ------------------------------------------------------------------- AXC
extern "C" int axspmv_build_axc(const axspmv_csr *a, const double *x, uint32_t hbs, uint32_t flags, axspmv_axc **out,
                                axspmv_fmtdata *fd) {
    AX_TRY(check_input(a, "build_axc"));
    if (!out || hbs == 0) return fail(AXSPMV_E_INVALID, "build_axc: hbs must be > 0");
    if ((flags & AXSPMV_BUILD_AX) && !x) return fail(AXSPMV_E_INVALID, "build_axc: AXSPMV_BUILD_AX needs x");
    *out = nullptr;
    const double t0 = now_s();
    const uint32_t n = a->nrows;
    auto *m = host_calloc<axspmv_axc>(1);
    if (!m) return AXSPMV_E_NOMEM;
    m->nrows = n;
    m->nnz = a->nnz;
    m->hbs = hbs;
    m->lenBRP = n + 1;
    m->brp = host_calloc<uint32_t>(size_t(n) + 1);
    if (!m->brp) {
        axspmv_axc_free(m);
        return AXSPMV_E_NOMEM;
    }
    uint64_t pos = 0;
    for (uint32_t r = 0; r < n; ++r) {
        pos += 2ull * hbs * ceil_div(rlen(a, r), hbs);
        if (pos >= (1ull << 32)) {
            axspmv_axc_free(m);
            return fail(AXSPMV_E_OVERFLOW, "build_axc: lenAX exceeds 32 bits");
        }
        m->brp[r + 1] = uint32_t(pos);
    }
    m->lenAX = uint32_t(pos);
    m->lenMAPX = m->lenAX / 2 + 8;
    m->mapx = host_calloc<uint32_t>(m->lenMAPX);
    if (flags & AXSPMV_BUILD_AX) m->ax = host_calloc<double>(m->lenAX);
    if (flags & AXSPMV_BUILD_PLANES) m->vplane = host_calloc<double>(m->lenAX / 2);
    if (!m->mapx || ((flags & AXSPMV_BUILD_AX) && !m->ax) || ((flags & AXSPMV_BUILD_PLANES) && !m->vplane)) {
        axspmv_axc_free(m);
        return AXSPMV_E_NOMEM;
    }
    if (flags & AXSPMV_BUILD_PAD_COL) std::fill(m->mapx, m->mapx + m->lenMAPX, AXSPMV_PAD_COL);
#pragma omp parallel for schedule(static)
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t b = a->row[r], l = rlen(a, r), s = m->brp[r] / 2;
        for (uint32_t k = 0; k < l; ++k) {
            const uint64_t q = uint64_t(s) + k;
            m->mapx[q] = a->col[b + k];
            if (m->vplane) m->vplane[q] = a->val[b + k];
            if (m->ax) {
                const uint64_t p = ax_of_plane(q, hbs);
                m->ax[p] = a->val[b + k];
                m->ax[p + hbs] = x[a->col[b + k]];
            }
        }
    }
    set_fd(fd, "faxc", 8.0 * m->lenAX + 4.0 * m->lenBRP, m->lenAX ? double(a->nnz) / double(m->lenAX / 2) : 0.0,
           now_s() - t0);
    *out = m;
    return AXSPMV_OK;
}

extern "C" void axspmv_axc_free(axspmv_axc *m) {
    if (!m) return;
    std::free(m->ax);
    std::free(m->brp);
    std::free(m->mapx);
    std::free(m->vplane);
    std::free(m);
}

// ------------------------------------------------------------------- AXT
extern "C" int axspmv_build_axt(const axspmv_csr *a, const double *x, uint32_t bs, uint32_t thw, uint32_t th,
                                uint32_t mode, uint32_t flags, axspmv_axt **out, axspmv_fmtdata *fd) {
    AX_TRY(check_input(a, "build_axt"));
    if (!out) return fail(AXSPMV_E_INVALID, "build_axt: out is NULL");
    *out = nullptr;
    if (thw == 0 || th == 0 || th > 1023) return fail(AXSPMV_E_INVALID, "build_axt: need thw > 0 and 0 < th < 1024");
    if (mode != AXSPMV_AXT_UNC && mode != AXSPMV_AXT_COM) return fail(AXSPMV_E_INVALID, "build_axt: mode must be UNC or COM");
    if ((flags & AXSPMV_BUILD_AX) && !x) return fail(AXSPMV_E_INVALID, "build_axt: AXSPMV_BUILD_AX needs x");
    const bool com = mode == AXSPMV_AXT_COM;
    uint32_t log = 0;
    if (com && th == 1) {
        // Header offsets are < bs and share the word with row << log(bs).
        if (!is_pow2(bs) || bs < 2 || bs > 1024)
            return fail(AXSPMV_E_INVALID, "build_axt: COM th==1 needs bs a power of two in [2,1024], got %u", bs);
        log = ilog2(bs);
    } else if (com) {
        if (!is_pow2(th)) return fail(AXSPMV_E_INVALID, "build_axt: COM needs th a power of two, got %u", th);
        log = ilog2(th);
    }
    if (com && a->nrows && (uint64_t(a->nrows - 1) << log) >= (1ull << 32))
        return fail(AXSPMV_E_OVERFLOW, "build_axt: row << %u overflows the 32-bit header for %u rows", log, a->nrows);
    const double t0 = now_s();
    const uint32_t n = a->nrows;
    const uint64_t ths = uint64_t(thw) * th;
    auto *m = host_calloc<axspmv_axt>(1);
    if (!m) return AXSPMV_E_NOMEM;
    m->nrows = n;
    m->nnz = a->nnz;
    m->mode = mode;
    m->tileHW = thw;
    m->tileH = th;
    m->logTH = ilog2(th);
    m->log = log;
    m->bs = bs;
    m->lenCON = n;
    m->con = host_calloc<uint32_t>(size_t(n) + 1);
    if (!m->con) {
        axspmv_axt_free(m);
        return AXSPMV_E_NOMEM;
    }
    // Lengths and row starts.
    uint64_t tileN = 0, lenSEC = 0;
    bool ovf = false;
    std::vector<uint64_t> colstart;  // UNC th>1: first global column of each row
    if (!com && th == 1) {
        uint64_t t = 0;
        for (uint32_t r = 0; r < n; ++r) {
            t += ceil_div(rlen(a, r), thw);
            if (t * 2 * thw >= (1ull << 32)) { ovf = true; break; }
            m->con[r + 1] = uint32_t(t * 2 * thw);
        }
        tileN = t;
        lenSEC = t;
    } else if (!com) {
        colstart.resize(size_t(n) + 1);
        uint64_t g = 0;
        for (uint32_t r = 0; r < n; ++r) {
            colstart[r] = g;
            const uint64_t tid = g / thw;
            // con[r] = first ax slot of the row (level 0 of its first column);
            // con[nrows] stays 0, as in the reference.
            const uint64_t c = tid * 2 * ths + (g - tid * thw);
            if (c >= (1ull << 32)) { ovf = true; break; }
            m->con[r] = uint32_t(c);
            g += ceil_div(rlen(a, r), th);
        }
        colstart[n] = g;
        tileN = ceil_div(g, thw);
        lenSEC = tileN * thw;
    } else {
        tileN = ceil_div(a->nnz, ths);
        lenSEC = tileN * ths;
    }
    const uint64_t lenAX = tileN * 2 * ths;
    if (ovf || lenAX >= (1ull << 32) || lenSEC >= (1ull << 32)) {
        axspmv_axt_free(m);
        return fail(AXSPMV_E_OVERFLOW, "build_axt: lenAX %llu exceeds 32 bits", (unsigned long long)lenAX);
    }
    m->tileN = uint32_t(tileN);
    m->lenAX = uint32_t(lenAX);
    m->lenSEC = uint32_t(lenSEC);
    m->sec = host_calloc<uint32_t>(lenSEC);
    if (flags & AXSPMV_BUILD_AX) m->ax = host_calloc<double>(lenAX);
    if (flags & AXSPMV_BUILD_PLANES) {
        m->vplane = host_calloc<double>(lenAX / 2);
        m->cplane = host_calloc<uint32_t>(lenAX / 2);
    }
    if (!m->sec || ((flags & AXSPMV_BUILD_AX) && !m->ax) ||
        ((flags & AXSPMV_BUILD_PLANES) && (!m->vplane || !m->cplane))) {
        axspmv_axt_free(m);
        return AXSPMV_E_NOMEM;
    }
    if ((flags & AXSPMV_BUILD_PAD_COL) && m->cplane) std::fill(m->cplane, m->cplane + lenAX / 2, AXSPMV_PAD_COL);
    auto put = [&](uint64_t q, uint32_t p) {
        if (m->vplane) {
            m->vplane[q] = a->val[p];
            m->cplane[q] = a->col[p];
        }
        if (m->ax) {
            const uint64_t s = ax_of_plane(q, thw);
            m->ax[s] = a->val[p];
            m->ax[s + thw] = x[a->col[p]];
        }
    };
#pragma omp parallel for schedule(dynamic, 256)
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t b = a->row[r], l = rlen(a, r);
        if (!com && th == 1) {
            const uint64_t s = m->con[r] / 2;  // plane slot of the row's first element
            for (uint32_t k = 0; k < l; ++k) {
                put(s + k, b + k);
                if (k % thw == 0) m->sec[(s + k) / thw] = r;
            }
        } else if (!com) {
            for (uint32_t k = 0; k < l; ++k) {
                const uint64_t g = colstart[r] + k / th, f = k % th;
                put((g / thw) * ths + f * thw + g % thw, b + k);
                if (f == 0) m->sec[g] = r;
            }
        } else if (th == 1) {
            for (uint32_t k = 0; k < l; ++k) {
                const uint32_t p = b + k, bco = p % bs;
                put(p, p);
                if (k == 0 || bco == 0) m->sec[p] = (r << log) | std::min(l - k - 1, bs - 1 - bco);
            }
        } else {
            for (uint32_t k = 0; k < l; ++k) {
                const uint64_t p = uint64_t(b) + k, t = p / ths, f = p % th, c = (p - t * ths) / th;
                const uint64_t q = t * ths + f * thw + c;
                put(q, uint32_t(p));
                if (k == 0 || f == 0) m->sec[q] = (r << log) | std::min<uint32_t>(l - k - 1, th - 1 - uint32_t(f));
            }
        }
    }
    char name[48];
    if (!com && th == 1)
        std::snprintf(name, sizeof name, "faxtuh1hw%u", thw);
    else if (!com)
        std::snprintf(name, sizeof name, "faxtuh%uhw%u", th, thw);
    else if (th == 1)
        std::snprintf(name, sizeof name, "faxtch1hw%ubs%u", thw, bs);
    else
        std::snprintf(name, sizeof name, "faxtch%uhw%u", th, thw);
    std::snprintf(m->name, sizeof m->name, "%s", name);
    set_fd(fd, name, 8.0 * double(lenAX) + 4.0 * double(lenSEC), lenAX ? double(a->nnz) / double(lenAX / 2) : 0.0,
           now_s() - t0);
    *out = m;
    return AXSPMV_OK;
}

extern "C" void axspmv_axt_free(axspmv_axt *m) {
    if (!m) return;
    std::free(m->ax);
    std::free(m->sec);
    std::free(m->con);
    std::free(m->vplane);
    std::free(m->cplane);
    std::free(m);
}
