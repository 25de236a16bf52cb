/*
 * ORACLE support -- not part of the library.
 *
 * main() for oracle/_ref/ref_dump_c<CHUNK>: the Makefile target `ref`
 * compiles the host-only line ranges of the reference's cudaSpmv.cu
 * (includes/macros, matrixReading, printMatrixStats, ncsr and the K1, AXC and
 * AXT builders) followed by this file, with g++, in one translation unit.
 * Nothing of the reference is copied into the repository; the binary lands
 * in oracle/_ref/ (git-ignored).  tests/golden/make_fixtures.py runs it to
 * write the fixture arrays the tests compare against, and the benchmark's
 * CPU baseline runs its `time` mode: the reference's own ncsr, repeated.
 *
 *   ref_dump M.bin X.bin OUT ncsr
 *   ref_dump M.bin X.bin OUT time SECONDS THREADS
 *   ref_dump M.bin X.bin OUT k1
 *   ref_dump M.bin X.bin OUT axc HBS
 *   ref_dump M.bin X.bin OUT axt BS THW TH UNC|COM
 *   ref_dump M.csr - OUT parse
 *
 * M.bin is read by the reference's matrixReading (.bin layout), X.bin holds
 * nrows doubles.  Each array is written raw to OUT.<name>; OUT.meta holds
 * "name value" lines (lengths, log, mfp, beta).  `parse` reads a matrix file
 * of either layout (.csr text or .bin) with matrixReading and writes the
 * arrays it parsed (row, col, val); it needs no x.
 */

static void dump_raw(const char *prefix, const char *name, const void *p, size_t bytes) {
    char path[512];
    snprintf(path, sizeof path, "%s.%s", prefix, name);
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); exit(2); }
    if (bytes && fwrite(p, 1, bytes, f) != bytes) { perror(path); exit(2); }
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s M.bin X.bin OUT ncsr|k1|axc HBS|axt BS THW TH MODE\n", argv[0]);
        return 2;
    }
    str_matCSR csr = matrixReading(argv[1]);
    if (!strcmp(argv[4], "parse")) {
        char metap[512];
        snprintf(metap, sizeof metap, "%s.meta", argv[3]);
        FILE *meta = fopen(metap, "w");
        if (!meta) { perror(metap); return 2; }
        fprintf(meta, "nrows %u\nnnz %u\n", csr.nrows, csr.nnz);
        fclose(meta);
        dump_raw(argv[3], "row", csr.row, (size_t)(csr.nrows + 1) * sizeof(UIN));
        dump_raw(argv[3], "col", csr.col, csr.nnz * sizeof(UIN));
        dump_raw(argv[3], "val", csr.val, csr.nnz * sizeof(FPT));
        return 0;
    }
    printMatrixStats(argv[1], &csr);
    FPT *x = (FPT *)calloc(csr.nrows + 1, sizeof(FPT));
    FILE *fx = fopen(argv[2], "rb");
    if (!fx || fread(x, sizeof(FPT), csr.nrows, fx) != csr.nrows) { fprintf(stderr, "bad x file\n"); return 2; }
    fclose(fx);
    const char *out = argv[3], *what = argv[4];
    char metap[512];
    snprintf(metap, sizeof metap, "%s.meta", out);
    FILE *meta = fopen(metap, "w");
    if (!meta) { perror(metap); return 2; }
    fprintf(meta, "nrows %u\nnnz %u\n", csr.nrows, csr.nnz);
    if (!strcmp(what, "time") && argc >= 7) {
        /* ncsr(THREADS) back to back until SECONDS have passed (the
         * reference's test_ncsr times single calls the same way) */
        const double limit = atof(argv[5]);
        const UIN nt = (UIN)atoi(argv[6]);
        FPT *y = (FPT *)calloc(csr.nrows + 1, sizeof(FPT));
        struct timespec t0, t1;
        unsigned long calls = 0;
        double el = 0.0;
        ncsr(nt, csr, x, y); /* first touch, untimed */
        clock_gettime(CLOCK_MONOTONIC, &t0);
        do {
            ncsr(nt, csr, x, y);
            ++calls;
            clock_gettime(CLOCK_MONOTONIC, &t1);
            el = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        } while (el < limit);
        dump_raw(out, "y", y, csr.nrows * sizeof(FPT));
        fprintf(meta, "threads %u\ncalls %lu\nseconds %.9f\n", nt, calls, el);
    } else if (!strcmp(what, "ncsr")) {
        FPT *y = (FPT *)calloc(csr.nrows + 1, sizeof(FPT));
        ncsr(1, csr, x, y);
        dump_raw(out, "y", y, csr.nrows * sizeof(FPT));
        dump_raw(out, "rl", csr.rl, csr.nrows * sizeof(UIN));
        fprintf(meta, "rmin %u\nrmax %u\nbw %u\nrave %.17g\nrsd %.17g\n", csr.rmin, csr.rmax, csr.bw,
                (double)csr.rave, (double)csr.rsd);
    } else if (!strcmp(what, "k1")) {
        str_matK1 k1;
        str_formatData fd = getFormatDataK1(CHUNK_SIZE, csr, x, &k1);
        fprintf(meta, "chunk %u\nchunkNum %u\nlenVC %u\nmfp %.17g\nbeta %.17g\n", (UIN)CHUNK_SIZE, k1.chunkNum,
                k1.lenVC, fd.mfp, fd.beta);
        dump_raw(out, "permi", k1.permi, (size_t)(k1.chunkNum + 1) * CHUNK_SIZE * sizeof(UIN));
        dump_raw(out, "nmc", k1.nmc, k1.chunkNum * sizeof(UIN));
        dump_raw(out, "chp", k1.chp, k1.chunkNum * sizeof(UIN));
        dump_raw(out, "val", k1.val, k1.lenVC * sizeof(FPT));
        dump_raw(out, "col", k1.col, k1.lenVC * sizeof(UIN));
    } else if (!strcmp(what, "axc") && argc >= 6) {
        str_matAXC axc;
        str_formatData fd = getFormatDataAXC(1, (UIN)atoi(argv[5]), csr, x, &axc);
        fprintf(meta, "hbs %u\nlenAX %u\nlenBRP %u\nlenMAPX %u\nmfp %.17g\nbeta %.17g\n", axc.hbs, axc.lenAX,
                axc.lenBRP, axc.lenMAPX, fd.mfp, fd.beta);
        dump_raw(out, "ax", axc.ax, axc.lenAX * sizeof(FPT));
        dump_raw(out, "brp", axc.brp, axc.lenBRP * sizeof(UIN));
        dump_raw(out, "mapx", axc.mapx, axc.lenMAPX * sizeof(UIN));
    } else if (!strcmp(what, "axt") && argc >= 9) {
        str_matAXT axt;
        memset(&axt, 0, sizeof axt);
        str_formatData fd = getFormatDataAXT(1, (UIN)atoi(argv[5]), (UIN)atoi(argv[6]), (UIN)atoi(argv[7]), argv[8],
                                             csr, x, &axt);
        fprintf(meta, "name %s\ntileN %u\nlenAX %u\nlenSEC %u\nlenCON %u\nlog %u\nlogTH %u\nmfp %.17g\nbeta %.17g\n",
                fd.name, axt.tileN, axt.lenAX, axt.lenSEC, axt.lenCON, axt.log, axt.logTH, fd.mfp, fd.beta);
        dump_raw(out, "ax", axt.ax, axt.lenAX * sizeof(FPT));
        dump_raw(out, "sec", axt.sec, axt.lenSEC * sizeof(UIN));
        dump_raw(out, "con", axt.con, (axt.lenCON + 1) * sizeof(UIN));
    } else {
        fprintf(stderr, "unknown request %s\n", what);
        return 2;
    }
    fclose(meta);
    return 0;
}
