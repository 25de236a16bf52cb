// Host CSR container, statistics, matrix files, transpose, RNG entry points.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <ctime>

#include "common.hpp"
#include "rng.hpp"

namespace axspmv {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
void clear_error() { g_err[0] = 0; }

double now_s() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return double(t.tv_sec) + 1e-9 * double(t.tv_nsec);
}

int check_csr_bounds(const axspmv_csr *a, const char *who) {
    if (!a || !a->row || (a->nnz && (!a->col || !a->val)))
        return fail(AXSPMV_E_INVALID, "%s: NULL matrix or array", who);
    if (a->row[0] != 0) return fail(AXSPMV_E_INVALID, "%s: row[0] = %u, expected 0", who, a->row[0]);
    if (a->row[a->nrows] != a->nnz)
        return fail(AXSPMV_E_INVALID, "%s: row[nrows] = %u but nnz = %u", who, a->row[a->nrows], a->nnz);
    for (uint32_t i = 0; i < a->nrows; ++i)
        if (a->row[i + 1] < a->row[i]) return fail(AXSPMV_E_INVALID, "%s: row offsets decrease at row %u", who, i);
    for (uint32_t j = 0; j < a->nnz; ++j)
        if (a->col[j] >= a->ncols)
            return fail(AXSPMV_E_INVALID, "%s: col[%u] = %u >= ncols %u", who, j, a->col[j], a->ncols);
    return AXSPMV_OK;
}

}  // namespace axspmv

using namespace axspmv;

extern "C" const char *axspmv_last_error(void) { return g_err; }
extern "C" int axspmv_abi_version(void) { return AXSPMV_ABI_VERSION; }

extern "C" int axspmv_csr_alloc(uint32_t nrows, uint32_t ncols, uint32_t nnz, axspmv_csr **out) {
    if (!out) return fail(AXSPMV_E_INVALID, "csr_alloc: out is NULL");
    *out = nullptr;
    auto *a = host_calloc<axspmv_csr>(1);
    if (!a) return AXSPMV_E_NOMEM;
    a->nrows = nrows;
    a->ncols = ncols;
    a->nnz = nnz;
    a->owner = 1;
    a->row = host_calloc<uint32_t>(size_t(nrows) + 1);
    a->col = host_calloc<uint32_t>(nnz);
    a->val = host_calloc<double>(nnz);
    a->rl = host_calloc<uint32_t>(nrows);
    if (!a->row || !a->col || !a->val || !a->rl) {
        axspmv_csr_free(a);
        return AXSPMV_E_NOMEM;
    }
    *out = a;
    return AXSPMV_OK;
}

extern "C" void axspmv_csr_free(axspmv_csr *a) {
    if (!a || !a->owner) return;
    std::free(a->row);
    std::free(a->col);
    std::free(a->val);
    std::free(a->rl);
    std::free(a);
}

extern "C" int axspmv_csr_check(const axspmv_csr *a) {
    if (!a || !a->row || (a->nnz && (!a->col || !a->val)))
        return fail(AXSPMV_E_INVALID, "csr_check: NULL matrix or array");
    if (a->row[0] != 0) return fail(AXSPMV_E_INVALID, "csr_check: row[0] = %u, expected 0", a->row[0]);
    if (a->row[a->nrows] != a->nnz)
        return fail(AXSPMV_E_INVALID, "csr_check: row[nrows] = %u but nnz = %u", a->row[a->nrows], a->nnz);
    for (uint32_t i = 0; i < a->nrows; ++i) {
        const uint32_t b = a->row[i], e = a->row[i + 1];
        if (e < b) return fail(AXSPMV_E_INVALID, "csr_check: row offsets decrease at row %u", i);
        for (uint32_t j = b; j < e; ++j) {
            if (a->col[j] >= a->ncols)
                return fail(AXSPMV_E_INVALID, "csr_check: col[%u] = %u >= ncols %u", j, a->col[j], a->ncols);
            if (j > b && a->col[j] <= a->col[j - 1])
                return fail(AXSPMV_E_INVALID, "csr_check: columns of row %u not strictly increasing", i);
        }
    }
    return AXSPMV_OK;
}

// printMatrixStats (cudaSpmv.cu:332-381) without the .sta write.  The
// reference forms rl*rl in 32-bit unsigned arithmetic (it wraps for rows
// longer than 65535); here the square is taken in double.
extern "C" int axspmv_csr_stats(axspmv_csr *a) {
    if (!a || !a->row || !a->rl) return fail(AXSPMV_E_INVALID, "csr_stats: NULL matrix or rl array");
    uint32_t rmin = 1000000000u, rmax = 0, bw = 0;
    double s1 = 0.0, s2 = 0.0;
    for (uint32_t i = 0; i < a->nrows; ++i) {
        const uint32_t rl = a->row[i + 1] - a->row[i];
        a->rl[i] = rl;
        s1 += rl;
        s2 += double(rl) * double(rl);
        rmin = std::min(rmin, rl);
        rmax = std::max(rmax, rl);
        for (uint32_t j = a->row[i]; j < a->row[i + 1]; ++j) {
            const int64_t d = int64_t(i) - int64_t(a->col[j]);
            const uint32_t ad = uint32_t(d < 0 ? -d : d);
            bw = std::max(bw, ad);
        }
    }
    const double n = a->nrows ? double(a->nrows) : 1.0;
    s1 /= n;
    s2 /= n;
    a->rmin = rmin;
    a->rmax = rmax;
    a->bw = bw;
    a->rave = s1;
    a->rsd = std::sqrt(std::max(0.0, s2 - s1 * s1));
    return AXSPMV_OK;
}

// The .sta layout of printMatrixStats (cudaSpmv.cu:359-379).
extern "C" int axspmv_csr_write_sta(const axspmv_csr *a, const char *matname, const char *path) {
    if (!a || !path) return fail(AXSPMV_E_INVALID, "csr_write_sta: NULL argument");
    FILE *fh = std::fopen(path, "w");
    if (!fh) return fail(AXSPMV_E_IO, "csr_write_sta: cannot open %s", path);
    const double rsdp = a->rave > 0 ? (a->rsd / a->rave) * 100.0 : 0.0;
    std::fprintf(fh, "------------------------------------\n");
    std::fprintf(fh, "matrix's statistics\n");
    std::fprintf(fh, "------------------------------------\n");
    std::fprintf(fh, "name:  %28s\n", matname ? matname : "");
    std::fprintf(fh, "nrows: %28u\n", a->nrows);
    std::fprintf(fh, "nnz:   %28u\n", a->nnz);
    std::fprintf(fh, "rmin:  %28u\n", a->rmin);
    std::fprintf(fh, "rave:  %28.2lf\n", a->rave);
    std::fprintf(fh, "rmax:  %28u\n", a->rmax);
    std::fprintf(fh, "rsd:   %28.2lf\n", a->rsd);
    std::fprintf(fh, "rsdp:  %28.2lf\n", rsdp);
    std::fprintf(fh, "bw:    %28u\n", a->bw);
    std::fclose(fh);
    return AXSPMV_OK;
}

static bool ends_with(const char *s, const char *suf) {
    const size_t n = std::strlen(s), m = std::strlen(suf);
    return n >= m && std::strcmp(s + n - m, suf) == 0;
}

// matrixReading (cudaSpmv.cu:265-328).  Unlike the reference, a short or
// malformed file is an error return, not an exit, and the matrix is checked.
extern "C" int axspmv_csr_read(const char *path, axspmv_csr **out) {
    if (!path || !out) return fail(AXSPMV_E_INVALID, "csr_read: NULL argument");
    *out = nullptr;
    const bool text = ends_with(path, ".csr"), bin = ends_with(path, ".bin");
    if (!text && !bin) return fail(AXSPMV_E_INVALID, "csr_read: %s has neither a .csr nor a .bin extension", path);
    FILE *fh = std::fopen(path, text ? "r" : "rb");
    if (!fh) return fail(AXSPMV_E_IO, "csr_read: unable to open %s", path);
    uint32_t nrows = 0, nnz = 0;
    axspmv_csr *a = nullptr;
    int rc = AXSPMV_OK;
    if (text) {
        if (std::fscanf(fh, "%u %u", &nrows, &nnz) != 2) {
            std::fclose(fh);
            return fail(AXSPMV_E_IO, "csr_read: bad header in %s", path);
        }
    } else {
        if (std::fread(&nrows, 4, 1, fh) != 1 || std::fread(&nnz, 4, 1, fh) != 1) {
            std::fclose(fh);
            return fail(AXSPMV_E_IO, "csr_read: short header in %s", path);
        }
    }
    rc = axspmv_csr_alloc(nrows, nrows, nnz, &a);
    if (rc) {
        std::fclose(fh);
        return rc;
    }
    if (text) {
        for (uint32_t i = 0; i < nnz && rc == AXSPMV_OK; ++i)
            if (std::fscanf(fh, "%lf %u", &a->val[i], &a->col[i]) != 2)
                rc = fail(AXSPMV_E_IO, "csr_read: entry %u of %s unreadable", i, path);
        for (uint32_t i = 0; i <= nrows && rc == AXSPMV_OK; ++i)
            if (std::fscanf(fh, "%u", &a->row[i]) != 1)
                rc = fail(AXSPMV_E_IO, "csr_read: row offset %u of %s unreadable", i, path);
    } else {
        if (std::fread(a->val, sizeof(double), nnz, fh) != nnz ||
            std::fread(a->col, 4, nnz, fh) != nnz ||
            std::fread(a->row, 4, size_t(nrows) + 1, fh) != size_t(nrows) + 1)
            rc = fail(AXSPMV_E_IO, "csr_read: %s is truncated", path);
    }
    std::fclose(fh);
    if (rc == AXSPMV_OK) {
        uint32_t maxc = 0;
        for (uint32_t j = 0; j < nnz; ++j) maxc = std::max(maxc, a->col[j] + 1);
        a->ncols = std::max(nrows, maxc);
        rc = axspmv_csr_check(a);
    }
    if (rc == AXSPMV_OK) rc = axspmv_csr_stats(a);
    if (rc) {
        axspmv_csr_free(a);
        return rc;
    }
    *out = a;
    return AXSPMV_OK;
}

extern "C" int axspmv_csr_write(const axspmv_csr *a, const char *path) {
    if (!a || !path) return fail(AXSPMV_E_INVALID, "csr_write: NULL argument");
    const bool text = ends_with(path, ".csr"), bin = ends_with(path, ".bin");
    if (!text && !bin) return fail(AXSPMV_E_INVALID, "csr_write: %s has neither a .csr nor a .bin extension", path);
    FILE *fh = std::fopen(path, text ? "w" : "wb");
    if (!fh) return fail(AXSPMV_E_IO, "csr_write: unable to open %s", path);
    bool ok = true;
    if (text) {
        ok = std::fprintf(fh, "%u %u\n", a->nrows, a->nnz) > 0;
        for (uint32_t i = 0; ok && i < a->nnz; ++i) ok = std::fprintf(fh, "%.17g %u\n", a->val[i], a->col[i]) > 0;
        for (uint32_t i = 0; ok && i <= a->nrows; ++i) ok = std::fprintf(fh, "%u\n", a->row[i]) > 0;
    } else {
        ok = std::fwrite(&a->nrows, 4, 1, fh) == 1 && std::fwrite(&a->nnz, 4, 1, fh) == 1 &&
             std::fwrite(a->val, sizeof(double), a->nnz, fh) == a->nnz &&
             std::fwrite(a->col, 4, a->nnz, fh) == a->nnz &&
             std::fwrite(a->row, 4, size_t(a->nrows) + 1, fh) == size_t(a->nrows) + 1;
    }
    if (std::fclose(fh) != 0) ok = false;
    return ok ? AXSPMV_OK : fail(AXSPMV_E_IO, "csr_write: writing %s failed", path);
}

// Counting-sort transpose: row j of A^T lists the rows i of A with A(i,j) != 0
// in increasing i, so columns of the result come out sorted.  perm (optional)
// receives, for every position q of A^T, the position in A it came from.
int axspmv::csr_transpose_perm(const axspmv_csr *a, axspmv_csr **out, std::vector<uint32_t> *perm) {
    if (!a || !out) return fail(AXSPMV_E_INVALID, "csr_transpose: NULL argument");
    *out = nullptr;
    AX_TRY(check_csr_bounds(a, "csr_transpose"));
    axspmv_csr *t = nullptr;
    AX_TRY(axspmv_csr_alloc(a->ncols, a->nrows, a->nnz, &t));
    for (uint32_t j = 0; j < a->nnz; ++j) t->row[a->col[j] + 1]++;
    for (uint32_t c = 0; c < a->ncols; ++c) t->row[c + 1] += t->row[c];
    std::vector<uint32_t> next(t->row, t->row + a->ncols);
    if (perm) perm->assign(a->nnz, 0);
    for (uint32_t i = 0; i < a->nrows; ++i)
        for (uint32_t j = a->row[i]; j < a->row[i + 1]; ++j) {
            const uint32_t q = next[a->col[j]]++;
            t->col[q] = i;
            t->val[q] = a->val[j];
            if (perm) (*perm)[q] = j;
        }
    int rc = axspmv_csr_stats(t);
    if (rc) {
        axspmv_csr_free(t);
        return rc;
    }
    *out = t;
    return AXSPMV_OK;
}

extern "C" int axspmv_csr_transpose(const axspmv_csr *a, axspmv_csr **out) {
    return csr_transpose_perm(a, out, nullptr);
}

extern "C" int axspmv_rng_u64(uint64_t seed, uint64_t stream, uint64_t *out, size_t n) {
    if (!out && n) return fail(AXSPMV_E_INVALID, "rng_u64: NULL output");
    Rng r(seed, stream);
    for (size_t i = 0; i < n; ++i) out[i] = r.next();
    return AXSPMV_OK;
}

extern "C" int axspmv_rng_uniform(uint64_t seed, uint64_t stream, double lo, double hi, double *out, size_t n) {
    if (!out && n) return fail(AXSPMV_E_INVALID, "rng_uniform: NULL output");
    Rng r(seed, stream);
    for (size_t i = 0; i < n; ++i) out[i] = r.uniform(lo, hi);
    return AXSPMV_OK;
}

extern "C" int axspmv_rng_normal(uint64_t seed, uint64_t stream, double mu, double sigma, double *out, size_t n) {
    if (!out && n) return fail(AXSPMV_E_INVALID, "rng_normal: NULL output");
    Rng r(seed, stream);
    for (size_t i = 0; i < n; ++i) out[i] = mu + sigma * r.normal();
    return AXSPMV_OK;
}
