/*
 * ORACLE -- test infrastructure, not part of the library (see oracle.h).
 * Compiled with -ffp-contract=off so every a + b*c is a separate multiply
 * and add, as in the reference's host build.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* cudaSpmv.cu:451-466 */
void orc_ncsr(int nthreads, uint32_t nrows, const uint32_t *row, const uint32_t *col,
              const double *val, const double *x, double *y) {
    long i;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (i = 0; i < (long)nrows; i++) {
        double aux = 0.0;
        for (uint32_t j = row[i]; j < row[i + 1]; j++) aux = aux + val[j] * x[col[j]];
        y[i] = aux;
    }
}

/* cudaSpmv.cu:451-466 with FPT = float for the values only. */
void orc_ncsr_f32val(int nthreads, uint32_t nrows, const uint32_t *row, const uint32_t *col,
                     const double *val, const double *x, double *y) {
    long i;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (i = 0; i < (long)nrows; i++) {
        double aux = 0.0;
        for (uint32_t j = row[i]; j < row[i + 1]; j++) aux = aux + (double)(float)val[j] * x[col[j]];
        y[i] = aux;
    }
}

void orc_spmm(int nthreads, uint32_t nrows, uint32_t k, const uint32_t *row, const uint32_t *col,
              const double *val, const double *X, double *Y) {
    long i;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (i = 0; i < (long)nrows; i++) {
        for (uint32_t c = 0; c < k; c++) {
            double aux = 0.0;
            for (uint32_t j = row[i]; j < row[i + 1]; j++) aux = aux + val[j] * X[(size_t)col[j] * k + c];
            Y[(size_t)i * k + c] = aux;
        }
    }
}

void orc_ncsr_t(uint32_t nrows, uint32_t ncols, const uint32_t *row, const uint32_t *col,
                const double *val, const double *ybar, double *xbar) {
    for (uint32_t c = 0; c < ncols; c++) xbar[c] = 0.0;
    for (uint32_t i = 0; i < nrows; i++)
        for (uint32_t j = row[i]; j < row[i + 1]; j++) xbar[col[j]] = xbar[col[j]] + val[j] * ybar[i];
}

void orc_vbar(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *ybar,
              const double *x, double *vbar) {
    for (uint32_t i = 0; i < nrows; i++)
        for (uint32_t j = row[i]; j < row[i + 1]; j++) vbar[j] = ybar[i] * x[col[j]];
}

/* cudaSpmv.cu:474-495 */
void orc_get_errors(uint32_t len, const double *ref, const double *res, int sign_blind,
                    double *aErr, double *rErr, uint32_t *pos) {
    double maxDif = 0.0, maxVal = 0.0;
    uint32_t p = 0;
    for (uint32_t i = 0; i < len; i++) {
        const double v = fabs(ref[i]);
        if (v > maxVal) maxVal = v;
        const double dif = sign_blind ? fabs(fabs(res[i]) - v) : fabs(res[i] - ref[i]);
        if (dif > maxDif) {
            maxDif = dif;
            p = i;
        }
    }
    *aErr = maxDif;
    *rErr = maxDif / maxVal; /* the reference divides unguarded: 0/0 = NaN */
    *pos = p;
}

/* cudaSpmv.cu:336-340 */
void orc_row_lengths(uint32_t nrows, const uint32_t *row, uint32_t *rl) {
    for (uint32_t i = 0; i < nrows; i++) rl[i] = row[i + 1] - row[i];
}

/* ------------------------------------------------------------------ K1 */
typedef struct { uint32_t ind, val; } pair_t;

/* orderFunction, cudaSpmv.cu:716-719 */
static int order_fn(const pair_t *a, const pair_t *b) { return (int)(b->val - a->val); }

/* Stable top-down merge sort: glibc 2.35's qsort is a merge sort (msort)
 * whenever its scratch buffer fits, which takes the left element on ties. */
static void msort_pairs(pair_t *a, pair_t *tmp, size_t n) {
    if (n < 2) return;
    const size_t n1 = n / 2, n2 = n - n1;
    msort_pairs(a, tmp, n1);
    msort_pairs(a + n1, tmp, n2);
    size_t i = 0, j = n1, k = 0;
    while (i < n1 && j < n) tmp[k++] = (order_fn(&a[i], &a[j]) <= 0) ? a[i++] : a[j++];
    while (i < n1) tmp[k++] = a[i++];
    memcpy(a, tmp, k * sizeof(pair_t));
}

/* getArrayPermiK1 + getArraysNmcChpK1, cudaSpmv.cu:723-754 */
uint32_t orc_k1_len(uint32_t nrows, const uint32_t *row, uint32_t chunk, int reorder,
                    uint32_t *permi, uint32_t *nmc, uint32_t *chp) {
    const uint32_t chunkNum = (nrows + chunk - 1) / chunk;
    pair_t *list = (pair_t *)malloc((nrows ? nrows : 1) * sizeof(pair_t));
    pair_t *tmp = (pair_t *)malloc((nrows ? nrows : 1) * sizeof(pair_t));
    for (uint32_t i = 0; i < nrows; i++) {
        list[i].ind = i;
        list[i].val = row[i + 1] - row[i];
    }
    if (reorder) msort_pairs(list, tmp, nrows);
    for (uint32_t i = 0; i < (chunkNum + 1) * chunk; i++) permi[i] = 0;
    for (uint32_t i = 0; i < nrows; i++) permi[i] = list[i].ind;
    uint32_t l = 0;
    for (uint32_t i = 0; i < chunkNum; i++) {
        uint32_t n = 0;
        if (reorder) {
            const uint32_t p = permi[i * chunk];
            n = row[p + 1] - row[p];
        } else {
            for (uint32_t r = i * chunk; r < (i + 1) * chunk && r < nrows; r++) {
                const uint32_t rl = row[permi[r] + 1] - row[permi[r]];
                if (rl > n) n = rl;
            }
        }
        nmc[i] = n;
        l = l + chunk * n;
    }
    if (chunkNum) chp[0] = 0;
    for (uint32_t i = 1; i < chunkNum; i++) chp[i] = chp[i - 1] + (nmc[i - 1] * chunk);
    free(list);
    free(tmp);
    return l;
}

/* getArraysValColK1, cudaSpmv.cu:758-778 */
void orc_k1_fill(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *val,
                 uint32_t chunk, const uint32_t *permi, const uint32_t *chp, double *kval,
                 uint32_t *kcol) {
    const uint32_t chunkNum = (nrows + chunk - 1) / chunk;
    for (uint32_t chunkID = 0; chunkID < chunkNum; chunkID++) {
        for (uint32_t rid = 0; rid < chunk; rid++) {
            uint32_t r = chunkID * chunk + rid;
            if (r == nrows) return;
            r = permi[r];
            uint32_t posCSR, rowOff;
            for (posCSR = row[r], rowOff = 0; posCSR < row[r + 1]; posCSR++, rowOff++) {
                const uint32_t posK1 = chp[chunkID] + rowOff * chunk + rid;
                kval[posK1] = val[posCSR];
                kcol[posK1] = col[posCSR];
            }
        }
    }
}

/* ----------------------------------------------------------------- AXC */
/* get_brpAXC, cudaSpmv.cu:937-948 */
uint32_t orc_axc_len(uint32_t nrows, const uint32_t *row, uint32_t hbs, uint32_t *brp) {
    brp[0] = 0;
    for (uint32_t r = 0; r < nrows; r++) {
        const uint32_t rl = row[r + 1] - row[r];
        const uint32_t brickNum = (rl + hbs - 1) / hbs;
        brp[r + 1] = brp[r] + (2 * brickNum * hbs);
    }
    return brp[nrows];
}

/* get_axAXC + get_mapxAXC, cudaSpmv.cu:952-1000 */
void orc_axc_fill(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *val,
                  const double *x, uint32_t hbs, const uint32_t *brp, double *ax, uint32_t *mapx) {
    for (uint32_t r = 0; r < nrows; r++) {
        uint32_t posAX = brp[r], counter = 0;
        for (uint32_t posCSR = row[r]; posCSR < row[r + 1]; posCSR++) {
            ax[posAX] = val[posCSR];
            ax[posAX + hbs] = x[col[posCSR]];
            if (counter == (hbs - 1)) {
                posAX = posAX + 1 + hbs;
                counter = 0;
            } else {
                posAX++;
                counter++;
            }
        }
    }
    for (uint32_t r = 0; r < nrows; r++) {
        uint32_t pos = brp[r] >> 1;
        for (uint32_t e = row[r]; e < row[r + 1]; e++) mapx[pos++] = col[e];
    }
}

/* ----------------------------------------------------------------- AXT */
void orc_axt_len(uint32_t nrows, const uint32_t *row, uint32_t thw, uint32_t th, uint32_t mode,
                 uint32_t *con, uint32_t *out3) {
    uint32_t tileN, lenAX, lenSEC;
    for (uint32_t r = 0; r <= nrows; r++) con[r] = 0;
    if (mode == 0 && th == 1) { /* getArraysLenAXT_UNC_H1, cudaSpmv.cu:1158-1175 */
        const uint32_t ts = 2 * thw;
        uint32_t totalTiles = 0;
        con[0] = 0;
        for (uint32_t r = 0; r < nrows; r++) {
            const uint32_t tiles = (row[r + 1] - row[r] + thw - 1) / thw;
            totalTiles = totalTiles + tiles;
            con[r + 1] = con[r] + tiles * ts;
        }
        tileN = totalTiles;
        lenAX = totalTiles * ts;
        lenSEC = totalTiles;
    } else if (mode == 0) { /* getArraysLenAXT_UNC, cudaSpmv.cu:1179-1201 */
        const uint32_t ths = thw * th;
        uint32_t rowStartPos = 0, totalColumns = 0;
        for (uint32_t r = 0; r < nrows; r++) {
            const uint32_t tid = ((rowStartPos + ths) / ths) - 1;
            const uint32_t fid = rowStartPos % th;
            const uint32_t cid = ((rowStartPos + th) / th) - 1 - (tid * thw);
            con[r] = tid * (2 * ths) + fid * (2 * thw) + cid;
            const uint32_t positions = (((row[r + 1] - row[r]) + th - 1) / th) * th;
            totalColumns = totalColumns + ((positions + th - 1) / th);
            rowStartPos = rowStartPos + positions;
        }
        tileN = (totalColumns + thw - 1) / thw;
        lenAX = tileN * 2 * ths;
        lenSEC = tileN * thw;
    } else { /* getArraysLenAXT_COM_H1 / _COM, cudaSpmv.cu:1205-1235 */
        const uint32_t ths = thw * th;
        const uint32_t totalElements = row[nrows];
        tileN = (totalElements + ths - 1) / ths;
        lenAX = 2 * tileN * ths;
        lenSEC = tileN * ths;
    }
    out3[0] = tileN;
    out3[1] = lenAX;
    out3[2] = lenSEC;
}

uint32_t orc_axt_fill(uint32_t nrows, const uint32_t *row, const uint32_t *col, const double *val,
                      const double *x, uint32_t bs, uint32_t thw, uint32_t th, uint32_t mode,
                      const uint32_t *con, double *ax, uint32_t *sec) {
    const uint32_t ths = thw * th;
    uint32_t log = 0;
    if (mode == 0 && th == 1) { /* getArraysAxSecAXT_UNC_H1, cudaSpmv.cu:1239-1271 */
        for (uint32_t r = 0; r < nrows; r++) {
            if (row[r + 1] - row[r] == 0) continue;
            uint32_t posAX = con[r], posSEC = posAX / (2 * ths), ctrEle = 0;
            for (uint32_t posCSR = row[r]; posCSR < row[r + 1]; posCSR++) {
                ax[posAX] = val[posCSR];
                ax[posAX + thw] = x[col[posCSR]];
                sec[posSEC] = r;
                posAX++;
                ctrEle++;
                if ((ctrEle % thw) == 0) {
                    posAX = posAX + thw;
                    posSEC++;
                }
            }
        }
    } else if (mode == 0) { /* getArraysAxSecAXT_UNC, cudaSpmv.cu:1275-1308 */
        for (uint32_t r = 0; r < nrows; r++) {
            if (row[r + 1] - row[r] == 0) continue;
            uint32_t posAX = con[r], posSEC = (posAX / (2 * ths)) * thw + posAX % thw, ctrEle = 0;
            for (uint32_t posCSR = row[r]; posCSR < row[r + 1]; posCSR++) {
                ax[posAX] = val[posCSR];
                ax[posAX + thw] = x[col[posCSR]];
                sec[posSEC] = r;
                posAX = posAX + 2 * thw;
                ctrEle++;
                if ((ctrEle % th) == 0) {
                    posAX = posAX + 1 - (2 * th * thw);
                    posSEC++;
                    if (posAX % thw == 0) posAX = posAX + ((2 * th) - 1) * thw;
                }
            }
        }
    } else if (th == 1) { /* getArraysAxSecAXT_COM_H1 + log, cudaSpmv.cu:1312-1345, 1504-1511 */
        for (uint32_t i = 1; i < 11; i++)
            if ((bs >> i) == 1) {
                log = i;
                break;
            }
        for (uint32_t r = 0; r < nrows; r++) {
            const uint32_t rowLen = row[r + 1] - row[r];
            if (rowLen == 0) continue;
            uint32_t eleCtr = 0;
            for (uint32_t posCSR = row[r]; posCSR < row[r + 1]; posCSR++) {
                const uint32_t bco = posCSR % bs;
                const uint32_t tid = ((posCSR + thw) / thw) - 1;
                const uint32_t tco = posCSR % thw;
                const uint32_t posAX = tid * 2 * thw + tco;
                const uint32_t posSEC = tid * thw + tco;
                ax[posAX] = val[posCSR];
                ax[posAX + thw] = x[col[posCSR]];
                if ((eleCtr == 0) || (bco == 0)) {
                    const uint32_t q1 = rowLen - eleCtr - 1, q2 = bs - 1 - bco;
                    sec[posSEC] = r << log | ((q1 > q2) ? q2 : q1);
                }
                eleCtr++;
            }
        }
    } else { /* getArraysAxSecAXT_COM + log, cudaSpmv.cu:1349-1387, 1529-1536 */
        for (uint32_t i = 1; i < 10; i++)
            if ((th >> i) == 1) {
                log = i;
                break;
            }
        const uint32_t ts = 2 * ths;
        for (uint32_t r = 0; r < nrows; r++) {
            const uint32_t rl = row[r + 1] - row[r];
            if (rl == 0) continue;
            uint32_t ec = 0;
            for (uint32_t pCSR = row[r]; pCSR < row[r + 1]; pCSR++) {
                const uint32_t tid = ((pCSR + ths) / ths) - 1;
                const uint32_t fid = pCSR % th;
                const uint32_t cid = (((pCSR - tid * ths) + th) / th) - 1;
                const uint32_t pAX = tid * ts + 2 * fid * thw + cid;
                const uint32_t pSEC = tid * ths + fid * thw + cid;
                ax[pAX] = val[pCSR];
                ax[pAX + thw] = x[col[pCSR]];
                if ((ec == 0) || (fid == 0)) {
                    const uint32_t q1 = rl - ec - 1, q2 = th - 1 - fid;
                    sec[pSEC] = r << log | ((q1 > q2) ? q2 : q1);
                }
                ec++;
            }
        }
    }
    return log;
}
