/* A C caller of libaxspmv, as the reference's C driver would bind it
 * (cudaSpmv.cu:2293-2495: read or make a matrix, build the formats, run and
 * time each kernel, compare with the CPU CSR product).  "host" mode touches
 * no GPU: generator, statistics, builders.  "gpu" mode also runs an AXT and
 * an RCT plan and checks y against a plain CSR loop. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "axspmv.h"

#define CHECK(call)                                                       \
    do {                                                                  \
        int rc_ = (call);                                                 \
        if (rc_ != AXSPMV_OK) {                                           \
            fprintf(stderr, "%s -> %d: %s\n", #call, rc_, axspmv_last_error()); \
            return 1;                                                     \
        }                                                                 \
    } while (0)

static void csr_product(const axspmv_csr *a, const double *x, double *y) {
    for (uint32_t r = 0; r < a->nrows; ++r) {
        double s = 0.0;
        for (uint32_t j = a->row[r]; j < a->row[r + 1]; ++j) s += a->val[j] * x[a->col[j]];
        y[r] = s;
    }
}

int main(int argc, char **argv) {
    const int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
    if (axspmv_abi_version() != AXSPMV_ABI_VERSION) return 2;
    axspmv_csr *a = NULL;
    CHECK(axspmv_gen_powerlaw(20000, 2.0, 512, 12.0, 7, &a));
    CHECK(axspmv_csr_stats(a));
    CHECK(axspmv_csr_check(a));
    double *x = malloc(sizeof(double) * a->ncols), *y = malloc(sizeof(double) * a->nrows),
           *ref = malloc(sizeof(double) * a->nrows);
    CHECK(axspmv_rng_uniform(1, 1ull << 40, -1.0, 1.0, x, a->ncols));
    axspmv_axt *t = NULL;
    axspmv_fmtdata fd;
    CHECK(axspmv_build_axt(a, x, 256, 4, 1, AXSPMV_AXT_UNC, AXSPMV_BUILD_AX | AXSPMV_BUILD_PLANES, &t, &fd));
    printf("host: %s n=%u nnz=%u rmax=%u %s beta=%.4f\n", "powerlaw", a->nrows, a->nnz, a->rmax, fd.name, fd.beta);
    axspmv_axt_free(t);
    csr_product(a, x, ref);
    if (gpu) {
        int ndev = 0;
        CHECK(axspmv_device_count(&ndev));
        if (ndev < 1) return 3;
        double *dx, *dy;
        CHECK(axspmv_dev_alloc(sizeof(double) * a->ncols, (void **)&dx));
        CHECK(axspmv_dev_alloc(sizeof(double) * a->nrows, (void **)&dy));
        CHECK(axspmv_memcpy_h2d(dx, x, sizeof(double) * a->ncols, NULL));
        const uint32_t formats[2] = {AXSPMV_FMT_AXT, AXSPMV_FMT_RCT};
        for (int f = 0; f < 2; ++f) {
            axspmv_opts o;
            CHECK(axspmv_opts_default(formats[f], &o));
            axspmv_plan *p = NULL;
            CHECK(axspmv_plan_create(a, &o, NULL, &p));
            axspmv_plan_info info;
            CHECK(axspmv_plan_get_info(p, &info));
            axspmv_timing tm;
            CHECK(axspmv_time_spmv(p, dx, dy, 20, 2, AXSPMV_TIME_EVENTS, 1, NULL, &tm));
            CHECK(axspmv_memcpy_d2h(y, dy, sizeof(double) * a->nrows, NULL));
            double num = 0.0, den = 0.0;
            for (uint32_t r = 0; r < a->nrows; ++r) {
                num = fmax(num, fabs(y[r] - ref[r]));
                den = fmax(den, fabs(ref[r]));
            }
            printf("gpu: %s %.2f us rel err %.3e\n", info.name, tm.mean_ms * 1e3, num / den);
            if (!(num / den <= 1e-12)) return 4;
            CHECK(axspmv_plan_destroy(p));
        }
        CHECK(axspmv_dev_free(dx));
        CHECK(axspmv_dev_free(dy));
    }
    free(x);
    free(y);
    free(ref);
    axspmv_csr_free(a);
    return 0;
}
