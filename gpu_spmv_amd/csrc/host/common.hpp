// Shared host-side helpers: per-thread last-error message and status macros.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "axspmv.h"

namespace axspmv {

// Records a formatted message as the calling thread's last error and returns
// `code`, so call sites read `return fail(AXSPMV_E_INVALID, "...")`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Checked host allocation; returns nullptr (and records the error) on failure.
template <class T>
T *host_calloc(size_t n) {
    if (n == 0) n = 1;
    T *p = static_cast<T *>(std::calloc(n, sizeof(T)));
    if (!p) fail(AXSPMV_E_NOMEM, "host allocation of %zu bytes failed", n * sizeof(T));
    return p;
}

inline bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }
inline uint32_t ilog2(uint32_t v) { uint32_t r = 0; while (v >>= 1) ++r; return r; }
inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Wall-clock seconds (CLOCK_MONOTONIC, as the reference's GT macro).
double now_s();

// What every kernel and builder needs of a host CSR to stay inside its
// arrays: non-NULL arrays, row[0] = 0, non-decreasing offsets, row[nrows] =
// nnz and every column < ncols.  Unsorted rows and repeated columns are
// accepted (the products sum every entry, as the reference's kernels do);
// axspmv_csr_check also requires strictly increasing columns.  Messages
// start with `who`.
int check_csr_bounds(const axspmv_csr *a, const char *who);

// axspmv_csr_transpose that also returns, for every position of A^T, the
// position in A its entry came from (perm may be NULL).
int csr_transpose_perm(const axspmv_csr *a, axspmv_csr **out, std::vector<uint32_t> *perm);

// Reverse Cuthill-McKee order of a square matrix (q[k] = row of A numbered
// k) and B = Q A Q^T with, for every entry of B, its position in A (pos may
// be NULL).
int rcm_order(const axspmv_csr *a, std::vector<uint32_t> &q);
int csr_permute_sym(const axspmv_csr *a, const std::vector<uint32_t> &q, axspmv_csr **out, std::vector<uint32_t> *pos);

}  // namespace axspmv

#define AX_TRY(expr)                        \
    do {                                    \
        int ax_rc_ = (expr);                \
        if (ax_rc_ != AXSPMV_OK) return ax_rc_; \
    } while (0)
