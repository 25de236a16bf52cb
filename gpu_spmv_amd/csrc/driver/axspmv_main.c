/*
 * axspmv_main -- the C host driver: reads or generates a matrix, builds
 * every format through the C ABI of libaxspmv.so, runs and times each GPU
 * kernel, checks it against the CPU CSR product, and prints the formats
 * table and the kernels table.
 *
 * It takes the place of the reference's main() (cudaSpmv.cu:2293-2495):
 * the same sequence (matrixReading + printMatrixStats, init_vec, ncsr and
 * test_ncsr, then each format's getFormatData* and test_g*) and the same two
 * tables ("formats' data": format, memory [Mbytes], occupancy, convTime [s];
 * "SpMV kernels' results": kernel, exeTime [s], Gflops, ordTime [s],
 * errAbs, errRel, rowInd; cudaSpmv.cu:2420-2492), with the wave64 formats'
 * parameters.  Differences: GPU times are per-launch hipEvent means over
 * NUM_ITE launches of kernels that write y completely (no y memset to
 * leave out); ordTime is the x refresh of the interleaved AX layout (the
 * other plans gather x inside the kernel, so 0); errors are printed both as
 * the reference computes them (get_errors compares |y| with |y_ref|,
 * cudaSpmv.cu:474-495) and signed (|y - y_ref|).
 *
 *   axspmv_main MATRIX [-b CBS] [-t THREADS] [-i NUM_ITE] [-x index|SEED] [-s] [-r]
 *
 * MATRIX is a .bin or .csr file (the reference's two layouts), or a
 * generator: laplace2d:GRID, powerlaw:N, stencil27:GRID (permuted, seed 2),
 * banded:N:HBW, random:N:RLMIN:RLMAX.  -b is the block size of the COM th =
 * 1 layout (the reference passes its CUDA block size there, default 256),
 * -t the OpenMP threads of the CPU product, -i NUM_ITE (default 250, as the
 * reference), -x the vector: "index" = x_i = i (the reference's init_vec),
 * otherwise U[-1, 1] from that seed (default 1), -s writes MATRIX's .sta
 * statistics file (matrix files only), -r adds rows for RCT and AXT UNC th
 * = 1 (thw 4) plans renumbered by reverse Cuthill-McKee inside the plan
 * (AXSPMV_PLAN_REORDER; square matrices; no reference counterpart).
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "axspmv.h"

#define CHECK(call)                                                                   \
    do {                                                                              \
        int rc_ = (call);                                                             \
        if (rc_ != AXSPMV_OK) {                                                       \
            fprintf(stderr, "axspmv_main: %s -> %d: %s\n", #call, rc_, axspmv_last_error()); \
            exit(EXIT_FAILURE);                                                       \
        }                                                                             \
    } while (0)

#define HDL printf("-----------------------------------------------------------------------------------------------------------\n")

typedef struct {
    double aerr, rerr; /* the reference's get_errors: max ||y_i| - |r_i||, / max |r_i| */
    int pos;
    double srerr;      /* signed: max |y_i - r_i| / max |r_i| */
} err_t;

typedef struct {
    char name[128];
    double et, gflops, ot;
    err_t e;
} row_t;

typedef struct {
    char name[80];
    double mfp, beta, ct, dct; /* dct: device conversion time (-1: none) */
} fmt_t;

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* y = A x on the host, OpenMP static over rows, left to right within a row */
static void cpu_csr(const axspmv_csr *a, const double *x, double *y, int threads) {
#pragma omp parallel for schedule(static) num_threads(threads)
    for (long r = 0; r < (long)a->nrows; ++r) {
        double s = 0.0;
        for (uint32_t j = a->row[r]; j < a->row[r + 1]; ++j) s += a->val[j] * x[a->col[j]];
        y[r] = s;
    }
}

static err_t errors(const double *ref, const double *y, uint32_t n) {
    err_t e = {0.0, 0.0, 0, 0.0};
    double mx = 0.0, sd = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double d = fabs(fabs(y[i]) - fabs(ref[i]));
        if (d > e.aerr) {
            e.aerr = d;
            e.pos = (int)i;
        }
        if (fabs(ref[i]) > mx) mx = fabs(ref[i]);
        if (fabs(y[i] - ref[i]) > sd) sd = fabs(y[i] - ref[i]);
    }
    e.rerr = mx > 0.0 ? e.aerr / mx : e.aerr;
    e.srerr = mx > 0.0 ? sd / mx : sd;
    return e;
}

static axspmv_csr *make_matrix(const char *spec) {
    axspmv_csr *a = NULL;
    unsigned long p1 = 0, p2 = 0, p3 = 0;
    if (sscanf(spec, "laplace2d:%lu", &p1) == 1)
        CHECK(axspmv_gen_laplace2d((uint32_t)p1, 4.0, -1.0, &a));
    else if (sscanf(spec, "powerlaw:%lu", &p1) == 1)
        CHECK(axspmv_gen_powerlaw((uint32_t)p1, 2.0, 2048, 20.0, 1, &a));
    else if (sscanf(spec, "stencil27:%lu", &p1) == 1)
        CHECK(axspmv_gen_stencil27((uint32_t)p1, 26.0, -1.0, 1, 2, &a));
    else if (sscanf(spec, "banded:%lu:%lu", &p1, &p2) == 2)
        CHECK(axspmv_gen_banded((uint32_t)p1, (uint32_t)p2, 0.3, 5, &a));
    else if (sscanf(spec, "random:%lu:%lu:%lu", &p1, &p2, &p3) == 3)
        CHECK(axspmv_gen_random((uint32_t)p1, (uint32_t)p2, (uint32_t)p3, 9, &a));
    else
        CHECK(axspmv_csr_read(spec, &a));
    return a;
}

typedef struct {
    const char *label;
    uint32_t format, kernel, mode, thw, th, bs, hbs, chunk, reorder, flags;
} kern_t;

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s MATRIX [-b CBS] [-t THREADS] [-i NUM_ITE] [-x index|SEED] [-s] [-r]\n", argv[0]);
        return EXIT_FAILURE;
    }
    const char *spec = argv[1];
    uint32_t cbs = 256, iters = 250;
    int threads = omp_get_max_threads(), xindex = 0, sta = 0, reorder = 0;
    uint64_t xseed = 1;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "-b") && i + 1 < argc)
            cbs = (uint32_t)atoi(argv[++i]);
        else if (!strcmp(argv[i], "-t") && i + 1 < argc)
            threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-i") && i + 1 < argc)
            iters = (uint32_t)atoi(argv[++i]);
        else if (!strcmp(argv[i], "-x") && i + 1 < argc) {
            ++i;
            if (!strcmp(argv[i], "index"))
                xindex = 1;
            else
                xseed = strtoull(argv[i], NULL, 10);
        } else if (!strcmp(argv[i], "-s"))
            sta = 1;
        else if (!strcmp(argv[i], "-r"))
            reorder = 1;
        else {
            fprintf(stderr, "axspmv_main: unknown argument %s\n", argv[i]);
            return EXIT_FAILURE;
        }
    }
    if (threads < 1 || iters < 1) {
        fprintf(stderr, "axspmv_main: threads and NUM_ITE must be positive\n");
        return EXIT_FAILURE;
    }
    int ndev = 0;
    CHECK(axspmv_device_count(&ndev));
    printf("matrix: %s\nCBS: %u  OMP threads: %d  NUM_ITE: %u  x: %s\nHIP devices: %d\n", spec, cbs, threads,
           iters, xindex ? "x_i = i" : "U[-1,1]", ndev);
    if (ndev < 1) {
        fprintf(stderr, "axspmv_main: no HIP device\n");
        return EXIT_FAILURE;
    }

    /* matrix, statistics, vectors, reference result ------------------- */
    axspmv_csr *a = make_matrix(spec);
    CHECK(axspmv_csr_stats(a));
    if (sta) {
        /* the reference names it after the matrix file up to its first '.'
         * (printMatrixStats, cudaSpmv.cu:359-362) */
        char path[512];
        snprintf(path, sizeof path, "%s", spec);
        char *dot = strchr(path, '.');
        if (dot) *dot = '\0';
        strncat(path, ".sta", sizeof path - strlen(path) - 1);
        CHECK(axspmv_csr_write_sta(a, spec, path));
    }
    printf("nrows %u  ncols %u  nnz %u  rmin %u  rave %.3f  rmax %u  rsd %.3f  bw %u\n", a->nrows, a->ncols, a->nnz,
           a->rmin, a->rave, a->rmax, a->rsd, a->bw);
    const uint32_t n = a->nrows, nc = a->ncols;
    double *x = malloc(sizeof(double) * (nc ? nc : 1)), *ref = malloc(sizeof(double) * (n ? n : 1)),
           *y = malloc(sizeof(double) * (n ? n : 1));
    if (!x || !ref || !y) return EXIT_FAILURE;
    if (xindex)
        for (uint32_t i = 0; i < nc; ++i) x[i] = (double)i;
    else
        CHECK(axspmv_rng_uniform(xseed, 1ull << 40, -1.0, 1.0, x, nc));
    cpu_csr(a, x, ref, threads);

    row_t rows[64];
    fmt_t fmts[32];
    int nr = 0, nf = 0;

    /* CPU CSR product, timed as test_ncsr (cudaSpmv.cu:503-527) */
    {
        double t = 0.0;
        for (uint32_t i = 0; i < iters; ++i) {
            const double t0 = now_s();
            cpu_csr(a, x, y, threads);
            t += now_s() - t0;
        }
        row_t *r = &rows[nr++];
        snprintf(r->name, sizeof r->name, "ncsr (%d threads)", threads);
        r->et = t / iters;
        r->gflops = 2.0 * a->nnz / r->et * 1e-9;
        r->ot = 0.0;
        r->e = errors(ref, y, n);
    }

    /* formats' data: the host builders, and the device builders' time ---- */
    {
        fmt_t *f = &fmts[nf++];
        snprintf(f->name, sizeof f->name, "fcsr");
        f->mfp = 12.0 * a->nnz + 4.0 * (n + 1) + 8.0 * n;
        f->beta = 1.0;
        f->ct = 0.0;
        f->dct = -1.0;
    }
    double *dx = NULL, *dy = NULL;
    CHECK(axspmv_dev_alloc(sizeof(double) * (nc ? nc : 1), (void **)&dx));
    CHECK(axspmv_dev_alloc(sizeof(double) * (n ? n : 1), (void **)&dy));
    CHECK(axspmv_memcpy_h2d(dx, x, sizeof(double) * nc, NULL));
    axspmv_dcsr dc = {n, nc, a->nnz, 0, NULL, NULL, NULL};
    CHECK(axspmv_dev_alloc(4ull * (n + 1), (void **)&dc.row));
    CHECK(axspmv_dev_alloc(4ull * (a->nnz ? a->nnz : 1), (void **)&dc.col));
    CHECK(axspmv_dev_alloc(8ull * (a->nnz ? a->nnz : 1), (void **)&dc.val));
    CHECK(axspmv_memcpy_h2d((void *)dc.row, a->row, 4ull * (n + 1), NULL));
    CHECK(axspmv_memcpy_h2d((void *)dc.col, a->col, 4ull * a->nnz, NULL));
    CHECK(axspmv_memcpy_h2d((void *)dc.val, a->val, 8ull * a->nnz, NULL));
    {
        axspmv_fmtdata fd;
        axspmv_k1 *k1 = NULL;
        CHECK(axspmv_build_k1(a, 64, 1, &k1, &fd));
        axspmv_k1_free(k1);
        fmt_t *f = &fmts[nf++];
        snprintf(f->name, sizeof f->name, "%s (C=64)", fd.name);
        f->mfp = fd.mfp, f->beta = fd.beta, f->ct = fd.ct, f->dct = -1.0;
        axspmv_axc *axc = NULL;
        CHECK(axspmv_build_axc(a, x, 64, AXSPMV_BUILD_AX, &axc, &fd));
        axspmv_axc_free(axc);
        f = &fmts[nf++];
        snprintf(f->name, sizeof f->name, "%s (hbs=64)", fd.name);
        f->mfp = fd.mfp, f->beta = fd.beta, f->ct = fd.ct;
        axspmv_dfmt *df = NULL;
        CHECK(axspmv_dev_build_axc(&dc, dx, 64, NULL, &df));
        f->dct = df->ms * 1e-3;
        axspmv_dfmt_free(df);
        /* the reference's AXT set (cudaSpmv.cu:2345-2356) at wave64 widths */
        static const uint32_t axt[][3] = {{AXSPMV_AXT_UNC, 16, 1}, {AXSPMV_AXT_UNC, 8, 1}, {AXSPMV_AXT_UNC, 4, 1},
                                          {AXSPMV_AXT_UNC, 64, 4}, {AXSPMV_AXT_UNC, 64, 6}, {AXSPMV_AXT_UNC, 64, 8},
                                          {AXSPMV_AXT_UNC, 64, 10}, {AXSPMV_AXT_UNC, 64, 12}, {AXSPMV_AXT_COM, 64, 1},
                                          {AXSPMV_AXT_COM, 64, 4}, {AXSPMV_AXT_COM, 64, 8}, {AXSPMV_AXT_COM, 64, 16}};
        for (size_t i = 0; i < sizeof axt / sizeof axt[0]; ++i) {
            axspmv_axt *t = NULL;
            CHECK(axspmv_build_axt(a, x, cbs, axt[i][1], axt[i][2], axt[i][0], AXSPMV_BUILD_AX, &t, &fd));
            axspmv_axt_free(t);
            f = &fmts[nf++];
            snprintf(f->name, sizeof f->name, "%s", fd.name);
            f->mfp = fd.mfp, f->beta = fd.beta, f->ct = fd.ct;
            CHECK(axspmv_dev_build_axt(&dc, dx, cbs, axt[i][1], axt[i][2], axt[i][0], NULL, &df));
            f->dct = df->ms * 1e-3;
            axspmv_dfmt_free(df);
        }
    }

    /* SpMV kernels ----------------------------------------------------- */
    static const kern_t ks[] = {
        {"gcsr (csr_naive)", AXSPMV_FMT_CSR, AXSPMV_KER_NAIVE, 0, 0, 0, 0, 0, 0, 0, 0},
        {"csr_vector", AXSPMV_FMT_CSR, AXSPMV_KER_STREAM, 0, 0, 0, 0, 0, 0, 0, 0},
        {"gk1 (k1_c64)", AXSPMV_FMT_K1, 0, 0, 0, 0, 0, 0, 64, 1, 0},
        {"gk1 (k1_c64_natural)", AXSPMV_FMT_K1, 0, 0, 0, 0, 0, 0, 64, 0, 0},
        {"gaxc (axc_hbs64)", AXSPMV_FMT_AXC, 0, 0, 0, 0, 0, 64, 0, 0, 0},
        {"gaxc (axc, auto hbs)", AXSPMV_FMT_AXC, 0, 0, 0, 0, 0, 0, 0, 0, 0},
        {"gaxtuh1hw16", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 16, 1, 0, 0, 0, 0, 0},
        {"gaxtuh1hw08", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 8, 1, 0, 0, 0, 0, 0},
        {"gaxtuh1hw04", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 4, 1, 0, 0, 0, 0, 0},
        {"gaxtuh1hw04 (ax layout)", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 4, 1, 0, 0, 0, 0, AXSPMV_PLAN_AX_LAYOUT},
        {"gaxtuh4hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 64, 4, 0, 0, 0, 0, 0},
        {"gaxtuh6hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 64, 6, 0, 0, 0, 0, 0},
        {"gaxtuh8hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 64, 8, 0, 0, 0, 0, 0},
        {"gaxtuh10hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 64, 10, 0, 0, 0, 0, 0},
        {"gaxtuh12hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 64, 12, 0, 0, 0, 0, 0},
        {"gaxtch1hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_COM, 64, 1, 1, 0, 0, 0, 0},
        {"gaxtch4hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_COM, 64, 4, 0, 0, 0, 0, 0},
        {"gaxtch8hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_COM, 64, 8, 0, 0, 0, 0, 0},
        {"gaxtch16hw64", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_COM, 64, 16, 0, 0, 0, 0, 0},
        {"rct (default tiles)", AXSPMV_FMT_RCT, 0, 0, 0, 0, 0, 0, 0, 0, 0},
        {"rct, renumbered", AXSPMV_FMT_RCT, 0, 0, 0, 0, 0, 0, 0, 0, AXSPMV_PLAN_REORDER},
        {"gaxtuh1hw04, renumbered", AXSPMV_FMT_AXT, 0, AXSPMV_AXT_UNC, 4, 1, 0, 0, 0, 0, AXSPMV_PLAN_REORDER},
    };
    for (size_t i = 0; i < sizeof ks / sizeof ks[0]; ++i) {
        const kern_t *k = &ks[i];
        if ((k->flags & AXSPMV_PLAN_REORDER) && !reorder) continue;
        axspmv_opts o;
        CHECK(axspmv_opts_default(k->format, &o));
        o.kernel = k->kernel;
        o.mode = k->mode;
        if (k->thw) o.thw = k->thw;
        if (k->th) o.th = k->th;
        if (k->bs) o.bs = cbs;
        if (k->format == AXSPMV_FMT_AXC || k->format == AXSPMV_FMT_CSR) o.hbs = k->hbs;
        if (k->chunk) o.chunk = k->chunk;
        if (k->format == AXSPMV_FMT_K1) o.reorder = k->reorder;
        o.flags = k->flags;
        axspmv_plan *p = NULL;
        CHECK(axspmv_plan_create(a, &o, NULL, &p));
        axspmv_plan_info info;
        CHECK(axspmv_plan_get_info(p, &info));
        axspmv_timing tm;
        CHECK(axspmv_time_spmv(p, dx, dy, iters, 5, AXSPMV_TIME_EVENTS, 1, NULL, &tm));
        row_t *r = &rows[nr++];
        snprintf(r->name, sizeof r->name, "%s [%s]", k->label, info.name);
        r->et = tm.mean_ms * 1e-3;
        r->gflops = 2.0 * a->nnz / r->et * 1e-9;
        r->ot = 0.0;
        if (k->flags & AXSPMV_PLAN_AX_LAYOUT) {
            axspmv_operands op;
            memset(&op, 0, sizeof op);
            op.x = dx;
            CHECK(axspmv_time_op(p, AXSPMV_OP_XREFRESH, &op, iters, 5, AXSPMV_TIME_EVENTS, 1, NULL, &tm));
            r->ot = tm.mean_ms * 1e-3;
        }
        CHECK(axspmv_spmv(p, dx, dy, NULL));
        CHECK(axspmv_memcpy_d2h(y, dy, sizeof(double) * n, NULL));
        r->e = errors(ref, y, n);
        CHECK(axspmv_plan_destroy(p));
        if (info.format == AXSPMV_FMT_RCT) {
            fmt_t *f = &fmts[nf++];
            snprintf(f->name, sizeof f->name, "%s", info.name);
            f->mfp = (double)info.device_bytes;
            f->beta = info.beta;
            f->ct = -1.0;
            f->dct = -1.0;
        }
    }

    HDL;
    printf("formats' data\n");
    HDL;
    printf("%25s %20s %10s %20s %20s\n", "format", "memory [Mbytes]", "occupancy", "convTime [s]", "devConvTime [s]");
    for (int i = 0; i < nf; ++i) {
        printf("%25s %20.2f %10.2f ", fmts[i].name, fmts[i].mfp * 1e-6, fmts[i].beta);
        if (fmts[i].ct >= 0.0)
            printf("%20.6f ", fmts[i].ct);
        else
            printf("%20s ", "-");
        if (fmts[i].dct >= 0.0)
            printf("%20.6f\n", fmts[i].dct);
        else
            printf("%20s\n", "-");
    }
    HDL;
    printf("SpMV kernels' results\n");
    HDL;
    printf("%40s %15s %8s %15s %13s %13s %10s %13s\n", "kernel", "exeTime [s]", "Gflops", "ordTime [s]", "errAbs",
           "errRel", "rowInd", "sErrRel");
    int bad = 0;
    for (int i = 0; i < nr; ++i) {
        printf("%40s %15.7f %8.3f %15.7f %13.3e %13.3e %10d %13.3e\n", rows[i].name, rows[i].et, rows[i].gflops,
               rows[i].ot, rows[i].e.aerr, rows[i].e.rerr, rows[i].e.pos, rows[i].e.srerr);
        if (!(rows[i].e.srerr <= 1e-12)) bad = 1;
    }
    fflush(stdout);
    axspmv_dev_free(dx);
    axspmv_dev_free(dy);
    axspmv_dev_free((void *)dc.row);
    axspmv_dev_free((void *)dc.col);
    axspmv_dev_free((void *)dc.val);
    free(x);
    free(y);
    free(ref);
    axspmv_csr_free(a);
    if (bad) {
        fprintf(stderr, "axspmv_main: a kernel's signed relative error exceeds 1e-12\n");
        return 4;
    }
    return EXIT_SUCCESS;
}
