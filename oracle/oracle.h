/*
 * ORACLE -- test infrastructure, not part of the library.
 *
 * A plain-C restatement of the CPU side of edoardoCB/gpu_spmv's SpMV path
 * (cudaSpmv.cu): the CSR product ncsr, the error metric get_errors, the row
 * statistics, and the K1 / AXC / AXT format builders, each following the
 * reference's own loops.  It is the checker the tests, the smoke test and
 * the benchmark's CPU-baseline leg compare the HIP library against; the
 * library never links or calls it.
 *
 * Every function names the reference lines it restates.  How each one is
 * pinned (fixtures produced by the compiled reference host code, closed
 * forms, brute force) is listed in tests/test_oracle.py.
 */
#ifndef AXSPMV_ORACLE_H
#define AXSPMV_ORACLE_H
#include <stdint.h>

/* ncsr, cudaSpmv.cu:451-466: y_i = sum_j val_j * x[col_j], left to right
 * from 0, one row per OpenMP iteration (schedule static). */
void orc_ncsr(int nthreads, uint32_t nrows, const uint32_t *row, const uint32_t *col,
              const double *val, const double *x, double *y);
/* ncsr with the values rounded to float first (FP32 value storage). */
void orc_ncsr_f32val(int nthreads, uint32_t nrows, const uint32_t *row, const uint32_t *col,
                     const double *val, const double *x, double *y);
/* k right-hand sides, X and Y interleaved (x[j*k + c]); per column c the
 * same left-to-right sum as ncsr. */
void orc_spmm(int nthreads, uint32_t nrows, uint32_t k, const uint32_t *row, const uint32_t *col,
              const double *val, const double *X, double *Y);
/* x_bar = A^T y_bar by scatter in row order (equals ncsr on A^T bit for bit). */
void orc_ncsr_t(uint32_t nrows, uint32_t ncols, const uint32_t *row, const uint32_t *col,
                const double *val, const double *ybar, double *xbar);
/* v_bar[k] = y_bar[row(k)] * x[col(k)] in CSR order. */
void orc_vbar(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *ybar,
              const double *x, double *vbar);

/* get_errors, cudaSpmv.cu:474-495.  sign_blind = 1 reproduces the
 * reference (| |ac| - |ar| |); 0 is the signed |ac - ar| the tests use. */
void orc_get_errors(uint32_t len, const double *ref, const double *res, int sign_blind,
                    double *aErr, double *rErr, uint32_t *pos);

/* rl[] of printMatrixStats, cudaSpmv.cu:336-340. */
void orc_row_lengths(uint32_t nrows, const uint32_t *row, uint32_t *rl);

/* K1, cudaSpmv.cu:716-846 (chunk = CHUNK_SIZE).  reorder = 1 sorts rows by
 * length with a stable merge sort (glibc's qsort); 0 keeps natural order with
 * nmc = longest row of the chunk.  Returns lenVC; permi has
 * (chunkNum+1)*chunk entries. */
uint32_t orc_k1_len(uint32_t nrows, const uint32_t *row, uint32_t chunk, int reorder,
                    uint32_t *permi, uint32_t *nmc, uint32_t *chp);
void orc_k1_fill(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *val,
                 uint32_t chunk, const uint32_t *permi, const uint32_t *chp, double *kval,
                 uint32_t *kcol);

/* AXC, cudaSpmv.cu:937-1000.  orc_axc_len fills brp and returns lenAX;
 * orc_axc_fill writes ax (lenAX) and mapx (lenAX/2 + 8). */
uint32_t orc_axc_len(uint32_t nrows, const uint32_t *row, uint32_t hbs, uint32_t *brp);
void orc_axc_fill(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *val,
                  const double *x, uint32_t hbs, const uint32_t *brp, double *ax, uint32_t *mapx);

/* AXT lengths, cudaSpmv.cu:1158-1235 (mode 0 = UNC, 1 = COM); con has
 * nrows+1 entries and is filled for UNC.  out3 = {tileN, lenAX, lenSEC}. */
void orc_axt_len(uint32_t nrows, const uint32_t *row, uint32_t thw, uint32_t th, uint32_t mode,
                 uint32_t *con, uint32_t *out3);
/* AXT fill, cudaSpmv.cu:1239-1387 (+ the log of getFormatDataAXT,
 * 1504-1511 / 1529-1536).  Returns log (0 for UNC). */
uint32_t orc_axt_fill(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *val,
                      const double *x, uint32_t bs, uint32_t thw, uint32_t th, uint32_t mode,
                      const uint32_t *con, double *ax, uint32_t *sec);
#endif
