/*
 * axspmv.h -- C ABI of the MI355X (gfx950) sparse matrix-vector library.
 *
 * The library reimplements the SpMV path of edoardoCB/gpu_spmv (file
 * cudaSpmv.cu) for one AMD Instinct MI355X: host format builders for CSR,
 * K1, AXC and AXT, and hand-written wave64 HIP kernels for y = A*x in each
 * format.  Every function returns an int status (AXSPMV_OK = 0, negative =
 * error) and never exits the process; axspmv_last_error() returns a message
 * for the calling thread's last failure.
 *
 * Conventions
 *  - Indices are uint32 (the reference's UIN), values are IEEE double unless
 *    a name says f32.  Matrices are square in the reference (x has nrows
 *    entries); ncols is kept for the transposed and rectangular cases.
 *  - Device pointers are plain pointers from hipMalloc (or from
 *    axspmv_dev_alloc); `stream` is a hipStream_t passed as void* (NULL =
 *    the default stream).  Device calls are asynchronous and stream ordered.
 *  - Host objects returned through `**out` are owned by the library and are
 *    released with the matching *_free().
 *  - A plan (device handle) is bound to the device that was current when it
 *    was created.  Distinct plans may be used from distinct threads; one plan
 *    must not be used by two threads at once.
 */
#ifndef AXSPMV_H
#define AXSPMV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXSPMV_ABI_VERSION 1

/* ---------------------------------------------------------------- status */
enum axspmv_status {
    AXSPMV_OK = 0,
    AXSPMV_E_INVALID = -1,     /* bad argument or inconsistent matrix      */
    AXSPMV_E_HIP = -2,         /* HIP runtime error                        */
    AXSPMV_E_NOMEM = -3,       /* host or device allocation failed         */
    AXSPMV_E_UNSUPPORTED = -4, /* parameter combination not implemented    */
    AXSPMV_E_OVERFLOW = -5,    /* 32-bit index or header packing overflow  */
    AXSPMV_E_IO = -6           /* matrix file could not be read / written  */
};

/* Message for the calling thread's most recent failure ("" if none).
 * Replaces the print-and-exit of HANDLE_CUDA_ERROR / TEST_POINTER / ABORT
 * (cudaSpmv.cu:146-148, 189-191, 255-257). */
const char *axspmv_last_error(void);
int axspmv_abi_version(void);

/* ------------------------------------------------------------ host CSR */
/* Host CSR matrix.  Mirrors str_matCSR (cudaSpmv.cu:261): val/col/row plus
 * the row-length array rl[] and the statistics printMatrixStats fills
 * (cudaSpmv.cu:332-381).  `owner` is 1 when the library allocated the
 * arrays (free with axspmv_csr_free), 0 for caller-owned arrays. */
typedef struct axspmv_csr {
    uint32_t nrows, ncols, nnz, owner;
    uint32_t *row;   /* nrows + 1 */
    uint32_t *col;   /* nnz, < ncols (any order within a row; see axspmv_csr_check) */
    double *val;     /* nnz */
    uint32_t *rl;    /* nrows (filled by axspmv_csr_stats) */
    uint32_t rmin, rmax, bw, pad_;
    double rave, rsd;
} axspmv_csr;

/* Allocate an empty library-owned CSR with the given sizes (arrays zeroed). */
int axspmv_csr_alloc(uint32_t nrows, uint32_t ncols, uint32_t nnz, axspmv_csr **out);
void axspmv_csr_free(axspmv_csr *a);

/* Structural checks: row[0]==0, row nondecreasing, row[nrows]==nnz,
 * col < ncols, columns strictly increasing within a row.  Plans, batches,
 * builders and the transpose check everything here but the column order
 * themselves (AXSPMV_E_INVALID before anything is built): unsorted rows and
 * repeated columns are accepted, every entry is summed, as by the
 * reference's kernels. */
int axspmv_csr_check(const axspmv_csr *a);

/* rl[i] = row[i+1]-row[i] and rmin/rave/rmax/rsd/bw.
 * Replaces printMatrixStats (cudaSpmv.cu:332-381) minus the file write. */
int axspmv_csr_stats(axspmv_csr *a);

/* The `.sta` statistics file printMatrixStats writes (cudaSpmv.cu:359-379). */
int axspmv_csr_write_sta(const axspmv_csr *a, const char *matname, const char *path);

/* Matrix files in the reference's two formats (matrixReading,
 * cudaSpmv.cu:265-328): ".csr" text "nrows nnz" then nnz lines "val col"
 * then nrows+1 row offsets; ".bin" = u32 nrows, u32 nnz, f64 val[nnz],
 * u32 col[nnz], u32 row[nrows+1].  The format follows the file extension. */
int axspmv_csr_read(const char *path, axspmv_csr **out);
int axspmv_csr_write(const axspmv_csr *a, const char *path);

/* A^T as a CSR (CSR -> CSC), columns sorted; used for the adjoint. */
int axspmv_csr_transpose(const axspmv_csr *a, axspmv_csr **out);

/* Reverse Cuthill-McKee order of a square matrix, on the pattern of
 * A + A^T: perm[k] (nrows entries) = the row/column of A numbered k.  Each
 * connected part starts from a pseudo-peripheral vertex (George-Liu
 * search from a vertex of least degree); neighbours are numbered by
 * increasing degree, ties by index; the order is then reversed.  The
 * ordering behind AXSPMV_PLAN_REORDER.  No reference counterpart (the
 * reference runs every matrix in the order read by matrixReading,
 * cudaSpmv.cu:265-328). */
int axspmv_csr_rcm(const axspmv_csr *a, uint32_t *perm);
/* B = P A P^T for a square A: row k of B is row perm[k] of A, column c of A
 * becomes the k with perm[k] = c, entries of a row sorted by the new
 * column.  perm must be a permutation of 0 .. nrows-1 (else
 * AXSPMV_E_INVALID).  Library-owned result (axspmv_csr_free). */
int axspmv_csr_permute(const axspmv_csr *a, const uint32_t *perm, axspmv_csr **out);

/* ------------------------------------------------------------ RNG + data */
/* xoshiro256** seeded through splitmix64(seed ^ stream * 0x9E3779B97F4A7C15).
 * u64 stream, U[lo,hi) doubles from (next>>11)*2^-53, and N(mu,sigma) by
 * Marsaglia's polar method. */
int axspmv_rng_u64(uint64_t seed, uint64_t stream, uint64_t *out, size_t n);
int axspmv_rng_uniform(uint64_t seed, uint64_t stream, double lo, double hi,
                       double *out, size_t n);
int axspmv_rng_normal(uint64_t seed, uint64_t stream, double mu, double sigma,
                      double *out, size_t n);

/* Synthetic matrices (no reference counterpart; the benchmark inputs).
 *  laplace2d : 5-point Laplacian on grid x grid, lexicographic order,
 *              diagonal `diag`, off-diagonals `off`.
 *  powerlaw  : n rows; z ~ Zipf(alpha) on {1..rlmax} by exact inverse CDF,
 *              rl = clamp(round(s*z), 1, rlmax) with s fixed so the mean row
 *              length is `mean`; columns uniform without replacement, sorted;
 *              values N(0,1).  Row i draws from stream i+1 of `seed`.
 *  stencil27 : 27-point stencil on grid^3, diagonal `diag`, off `off`; with
 *              permute != 0 the symmetric permutation P A P^T of a
 *              Fisher-Yates shuffle (seed `seed`, stream 0) is applied.
 *  banded    : n rows, every (i,j) with |i-j| <= hbw is kept with
 *              probability `fill` (the diagonal too), values N(0,1).
 *  random    : n rows, rl uniform in [rlmin, rlmax], distinct uniform
 *              columns, values N(0,1). */
int axspmv_gen_laplace2d(uint32_t grid, double diag, double off, axspmv_csr **out);
int axspmv_gen_powerlaw(uint32_t n, double alpha, uint32_t rlmax, double mean,
                        uint64_t seed, axspmv_csr **out);
int axspmv_gen_stencil27(uint32_t grid, double diag, double off, int permute,
                         uint64_t seed, axspmv_csr **out);
int axspmv_gen_banded(uint32_t n, uint32_t hbw, double fill, uint64_t seed,
                      axspmv_csr **out);
int axspmv_gen_random(uint32_t n, uint32_t rlmin, uint32_t rlmax, uint64_t seed,
                      axspmv_csr **out);
/* Overwrite the values of `a` with N(mu, sigma) draws (seed, stream 0). */
int axspmv_fill_values_normal(axspmv_csr *a, double mu, double sigma, uint64_t seed);

/* ------------------------------------------------------- format builders */
/* str_formatData (cudaSpmv.cu:385): name, memory footprint in bytes,
 * occupancy beta = nnz / value slots, and the build time in seconds. */
typedef struct axspmv_fmtdata {
    char name[48];
    double mfp, beta, ct;
} axspmv_fmtdata;

/* Builder flags. */
#define AXSPMV_BUILD_AX 1u     /* emit the reference's interleaved ax[] (needs x) */
#define AXSPMV_BUILD_PLANES 2u /* emit the value plane / column plane        */
#define AXSPMV_BUILD_PAD_COL 4u /* padding slots of the column plane (AXT) and
                                   of mapx (AXC) hold AXSPMV_PAD_COL instead
                                   of the reference's 0.  The device plans are
                                   built this way: their kernels read 0.0 for
                                   such a slot, so a non-finite x[0] does not
                                   reach rows that have padding (0 * inf).  */
#define AXSPMV_PAD_COL 0xffffffffu

/* K1 (str_matK1, cudaSpmv.cu:712; builders cudaSpmv.cu:716-846).
 * chunk = rows per chunk (the reference's CHUNK_SIZE; 64 = one wavefront).
 * reorder = 1: rows sorted by length, descending, ties by ascending row
 * (a stable sort; the reference's qsort is glibc's merge sort);
 * reorder = 0: natural row order, nmc[c] = longest row of the chunk. */
typedef struct axspmv_k1 {
    uint32_t nrows, nnz, chunk, chunkNum, lenVC, reorder;
    uint32_t *permi;  /* (chunkNum+1)*chunk, zero padded */
    uint32_t *nmc;    /* chunkNum */
    uint32_t *chp;    /* chunkNum */
    double *val;      /* lenVC */
    uint32_t *col;    /* lenVC */
} axspmv_k1;
int axspmv_build_k1(const axspmv_csr *a, uint32_t chunk, int reorder,
                    axspmv_k1 **out, axspmv_fmtdata *fd);
void axspmv_k1_free(axspmv_k1 *m);

/* AXC (str_matAXC, cudaSpmv.cu:933; builders cudaSpmv.cu:937-1062).
 * hbs = brick half size (the reference's TILE_HW = 32).  With
 * AXSPMV_BUILD_AX, ax[] holds bricks of hbs values followed by hbs gathered
 * x-values (x required); with AXSPMV_BUILD_PLANES, vplane[] holds the value
 * halves (lenAX/2) and mapx[] is the column plane. */
typedef struct axspmv_axc {
    uint32_t nrows, nnz, hbs, lenAX, lenBRP, lenMAPX;
    double *ax;       /* lenAX (AXSPMV_BUILD_AX) */
    uint32_t *brp;    /* nrows + 1 */
    uint32_t *mapx;   /* lenMAPX = lenAX/2 + 8 */
    double *vplane;   /* lenAX/2 (AXSPMV_BUILD_PLANES) */
} axspmv_axc;
int axspmv_build_axc(const axspmv_csr *a, const double *x, uint32_t hbs, uint32_t flags,
                     axspmv_axc **out, axspmv_fmtdata *fd);
void axspmv_axc_free(axspmv_axc *m);

/* AXT (str_matAXT, cudaSpmv.cu:1154; builders cudaSpmv.cu:1158-1566).
 * mode AXSPMV_AXT_UNC: uncompacted, tiles padded per row;
 * mode AXSPMV_AXT_COM: compacted, packed by nonzero index with segment
 * headers (row << log) | (segment length - 1).
 * th = tile height (1 = the "H1" layouts), thw = tile half width,
 * bs = block size (only COM with th == 1 uses it: headers split at bs).
 * vplane[]/cplane[] (lenAX/2 each) are the value halves of ax[] and the
 * column of every slot (0 in padding): x[cplane[q]] reproduces the x half. */
#define AXSPMV_AXT_UNC 0u
#define AXSPMV_AXT_COM 1u
typedef struct axspmv_axt {
    char name[48];
    uint32_t nrows, nnz, mode, tileHW, tileH, logTH, tileN;
    uint32_t lenAX, lenSEC, lenCON, log, bs;
    double *ax;        /* lenAX (AXSPMV_BUILD_AX) */
    uint32_t *sec;     /* lenSEC: row per tile (UNC H1), row per column
                          (UNC H>1) or per-slot headers (COM) */
    uint32_t *con;     /* lenCON + 1: first ax slot of each row (UNC) */
    double *vplane;    /* lenAX/2 (AXSPMV_BUILD_PLANES) */
    uint32_t *cplane;  /* lenAX/2 (AXSPMV_BUILD_PLANES) */
} axspmv_axt;
int axspmv_build_axt(const axspmv_csr *a, const double *x, uint32_t bs, uint32_t thw,
                     uint32_t th, uint32_t mode, uint32_t flags, axspmv_axt **out,
                     axspmv_fmtdata *fd);
void axspmv_axt_free(axspmv_axt *m);

/* ------------------------------------------------------- device plans */
/* Formats and kernels.  A plan uploads one matrix in one format and runs the
 * kernel chosen by (format, kernel). */
enum axspmv_format {
    AXSPMV_FMT_CSR = 0, /* val/col/row                                  */
    AXSPMV_FMT_K1 = 1,  /* sliced ELL, chunk rows per wavefront          */
    AXSPMV_FMT_AXC = 2, /* row bricks of hbs slots                       */
    AXSPMV_FMT_AXT = 3, /* tiled: UNC/COM x th==1 / th>1                 */
    AXSPMV_FMT_RCT = 4  /* row-block x column-slice tiles, entries sorted
                           by column, one workgroup per tile (no reference
                           counterpart: built for gather-bound matrices)  */
};
enum axspmv_kernel {
    AXSPMV_KER_DEFAULT = 0, /* the format's own kernel                  */
    AXSPMV_KER_NAIVE = 1,   /* CSR: one lane per row (gcsr)              */
    AXSPMV_KER_STREAM = 2   /* CSR: a group of hbs lanes per row (CSR
                               vector; hbs = 0 picks it from the mean row
                               length), rows over 256 entries one wave each */
};
enum axspmv_vtype { AXSPMV_F64 = 0, AXSPMV_F32 = 1 };

typedef struct axspmv_opts {
    uint32_t format;  /* enum axspmv_format */
    uint32_t kernel;  /* enum axspmv_kernel */
    uint32_t mode;    /* AXT: AXSPMV_AXT_UNC / AXSPMV_AXT_COM */
    uint32_t thw;     /* AXT tile half width (0 = 64, or 4 for UNC th==1) */
    uint32_t th;      /* AXT tile height (0 = 1) */
    uint32_t bs;      /* AXT COM th==1 block size (0 = 256) */
    uint32_t hbs;     /* AXC brick size; CSR vector kernel: lanes per row.
                         A power of two <= 64; 0 (the CSR and AXC default) =
                         the power of two g in [2, 64] with 1.5 g >= the
                         mean row length */
    uint32_t chunk;   /* K1 rows per chunk (0 = 64) */
    uint32_t reorder; /* K1: 1 = sort rows by length (default), 0 = natural */
    uint32_t vtype;   /* enum axspmv_vtype: value storage */
    uint32_t block;   /* threads per workgroup (0 = kernel default; the
                         RCT default options keep 0: for SpMV 512, or 256 /
                         128 for tiles of under 16K / 2K entries; 1024 for
                         SpMM) */
    uint32_t budget;  /* units per wave block for the AXT UNC row-block kernels
                         (0 = 64, one pass of the wave; batches: 256) */
    uint32_t flags;   /* AXSPMV_PLAN_* */
    uint32_t ktile;   /* SpMM k-tile: right-hand sides per lane group, 8/16/32/64
                         (0 = smallest of those >= min(k, 64)) */
    uint32_t tile_rows; /* RCT: at most this many rows per row block (0 =
                           min(16384, LDS / 8)); blocks are cut to about
                           equal nonzeros */
    uint32_t slices;    /* RCT: column slices 1, 2, 4 or 8 (0 = one slice if
                           x fits 2 MB or the row blocks read column bands
                           of <= 2 MB of x -- banded matrices; else the
                           fewest slices for which the row blocks of full
                           height (16384 rows, or tile_rows) times the
                           slices cover the CUs, at most as many as give x
                           slices of <= 2 MB and at least as many as give
                           x slices of <= 8 MiB; the plan of A^T: half the
                           2 MB count).
                           AXT UNC th = 1: 2, 4 or 8 splits the matrix into
                           that many column slices, each stored in its own
                           AXT layout (the builders' arrays for that part of
                           the matrix), each XCD gathering from one slice of
                           x; y is cleared, then the slices' row sums are
                           added with FP64 atomics (rounding order varies
                           between runs).  SpMV and the adjoint only; 0 or 1
                           = one layout; other AXT layouts reject S > 1. */
} axspmv_opts;

#define AXSPMV_PLAN_TRANSPOSE 1u /* also build A^T (needed by the adjoint) */
#define AXSPMV_PLAN_RCT_SPLIT 2u /* RCT: keep column (4 B) and tile row (2 B) of
                                    each entry apart even when both fit one
                                    32-bit key */
#define AXSPMV_PLAN_STAGE_X 8u   /* AXT UNC (plans and batches): when at least
                                    half of the wave blocks touch <= 1024
                                    consecutive columns (banded / stencil
                                    structure), each such block first copies
                                    its part of x to LDS and gathers from
                                    there.  Off by default: on gfx950 the
                                    vector L1 already serves these gathers and
                                    the staging step measured slower. */
#define AXSPMV_PLAN_DETERMINISTIC 16u /* results independent of the order in
                                    which workgroups run: AXT plans in column
                                    slices store each slice's row sums in a
                                    partial vector (slices x nrows doubles,
                                    held by the plan) and a second kernel adds
                                    them in slice order, instead of FP64
                                    atomics into y; RCT plans with
                                    AXSPMV_PLAN_TRANSPOSE also build A^T as
                                    such an AXT plan (two slots per tile,
                                    slices of <= 2 MB of y_bar) and
                                    axspmv_adjoint runs on it.  The RCT
                                    forward SpMV already adds its slices in
                                    slice order (partials and a merge), but
                                    its tiles sum their rows in LDS in no
                                    fixed order; its fused forward + adjoint
                                    keeps those LDS adds and FP64 slice adds
                                    into x_bar. */
#define AXSPMV_PLAN_RCT_ATOMIC 32u /* RCT with column slices: always clear y
                                    and add each slice's row sums into it
                                    with FP64 atomics.  By default a plan
                                    whose slices write many row sums
                                    against its nonzeros (slices x nrows >=
                                    nnz / 6) stores them in slice partials
                                    (slices x nrows doubles held by the
                                    plan) and a second kernel adds those in
                                    slice order; other plans add. */
#define AXSPMV_PLAN_REORDER 64u /* square matrices: build the plan for
                                    B = Q A Q^T, Q the reverse Cuthill-McKee
                                    order of axspmv_csr_rcm, and run every
                                    operation in the caller's numbering:
                                    x (and y_bar, b) gathered into Q's
                                    order, the format's kernels on B, the
                                    results gathered back.  x and y stay in
                                    A's numbering; v_bar comes out in the
                                    order of B's plan, and
                                    axspmv_plan_vbar_positions gives A's
                                    CSR positions.  For matrices whose
                                    numbering scatters a banded structure
                                    (a permuted stencil): B reads x through
                                    a narrow window.  Where B's entries lie
                                    on average not even twice as close to
                                    the diagonal as A's (random or already
                                    banded matrices) the plan is built for A
                                    as without the flag (its name then has
                                    no "_rcm").  The plan holds B and
                                    three (with AXSPMV_PLAN_TRANSPOSE six)
                                    vectors of nrows doubles; SpMM moves
                                    X's and Y's rows through 2 x nrows x k
                                    doubles the plan allocates at the first
                                    call with that k or more (a smaller
                                    earlier buffer is kept until the plan
                                    is destroyed, for graphs that use it).
                                    Every operation is available.  Building
                                    Q costs a few breadth-first passes over
                                    A on the host. */
#define AXSPMV_PLAN_RCT_WIDE_KEY 128u /* RCT: keep the 32-bit key where a
                                    16-bit one would fit.  By default a
                                    matrix whose entries lie close to the
                                    diagonal (a band of up to about 127 on
                                    each side) stores each entry's column and
                                    row in 16 bits -- the column's offset
                                    from the smallest column of its run of
                                    128 entries and the row minus the column
                                    -- and its plan's name ends in "_k16". */
#define AXSPMV_PLAN_AX_LAYOUT 4u /* AXT UNC th == 1 or AXC, FP64 values: keep
                                    the reference's interleaved ax array on the
                                    device (tiles of thw values then thw
                                    x-values, getArraysAxSecAXT_UNC_H1,
                                    cudaSpmv.cu:1239-1271; AXC's bricks,
                                    get_axAXC 952-979, are that layout with
                                    thw = hbs).  Every SpMV first
                                    refreshes the x-halves from x
                                    (axspmv_x_refresh), then streams ax like
                                    gaxtuh1hw*w (cudaSpmv.cu:1570-1850). */

/* Fill `o` with the defaults of a format (every field defined). */
int axspmv_opts_default(uint32_t format, axspmv_opts *o);

typedef struct axspmv_plan axspmv_plan;

/* Build (host) and upload (device) A in the format of `o`.  Replaces the
 * getFormatData*() + cudaMalloc/cudaMemcpy part of every test_g*()
 * (cudaSpmv.cu:547-563, 871-897, 1100-1119, 1603-1622, ...). */
int axspmv_plan_create(const axspmv_csr *a, const axspmv_opts *o, void *stream,
                       axspmv_plan **out);
int axspmv_plan_destroy(axspmv_plan *p);

typedef struct axspmv_plan_info {
    char name[64];            /* kernel name, e.g. "axt_unc_h1_thw4"      */
    uint32_t nrows, ncols, nnz, format;
    uint64_t slots;           /* value slots incl. padding                */
    uint64_t device_bytes;    /* bytes of matrix data held on the device  */
    uint64_t stream_bytes;    /* matrix bytes one SpMV reads (incl. pad)  */
    uint64_t alg_bytes;       /* algorithmic bytes of one FP64 SpMV:
                                 (s_v+4)*nnz + 4*(n+1) + 8*ncols + 8*nrows */
    double beta;              /* nnz / slots                              */
    uint32_t launches;        /* kernel launches per SpMV                 */
    uint32_t nblocks;         /* wave blocks (segmented kernels)          */
    uint32_t staged_blocks;   /* wave blocks that read x from LDS         */
    uint32_t pad_;
} axspmv_plan_info;
int axspmv_plan_get_info(const axspmv_plan *p, axspmv_plan_info *info);

/* y = A x.  y is fully overwritten (no zeroing needed by the caller).
 * Some plans keep intermediate sums in device buffers they own and use in
 * every call: RCT and AXSPMV_PLAN_DETERMINISTIC AXT plans with column
 * slices (slice partials), K1 and AXT COM (rows split between waves), the
 * AX layout (its refreshed x half).  Calls on one plan that overlap in time
 * (several streams or host threads) race; serialize them, or give each
 * stream its own plan. */
int axspmv_spmv(axspmv_plan *p, const double *d_x, double *d_y, void *stream);

/* ------------------------------------------------- adjoint (gradients) */
/* x_bar = A^T y_bar, run by the same kernel on a precomputed A^T in the same
 * format (plan created with AXSPMV_PLAN_TRANSPOSE).  CSR, K1, AXC, AXT COM
 * and AXT UNC plans without column slices reduce every row inside one wave
 * or a fixed fix-up order: no atomics, bit-identical results from call to
 * call.  RCT plans run the RCT plan of A^T, whose tiles add into LDS with
 * atomics and, with several column slices, into x_bar with FP64 atomics, and
 * AXT plans in column slices add slice sums with FP64 atomics: the rounding
 * of x_bar then varies between calls (within the FP64 tolerance).  With
 * AXSPMV_PLAN_DETERMINISTIC both are atomic-free: RCT plans run a
 * deterministic AXT plan of A^T in column slices, whose slice partials are
 * summed in slice order.  No reference counterpart (the reference computes
 * forward products only). */
int axspmv_adjoint(axspmv_plan *p, const double *d_ybar, double *d_xbar, void *stream);
/* Per-nonzero value gradient v_bar[k] = y_bar[row(k)] * x[col(k)].  Plans:
 * CSR naive and AXT UNC th == 1 write it in the CSR order of the matrix the
 * plan was built from; RCT (needs AXSPMV_PLAN_TRANSPOSE) writes it in the
 * entry order of its plan of A^T, and axspmv_plan_vbar_positions gives each
 * entry's CSR position. */
int axspmv_vbar(axspmv_plan *p, const double *d_ybar, const double *d_x, double *d_vbar, void *stream);
/* h_pos[e] (count = nnz entries) = CSR position of v_bar entry e: the
 * identity for CSR and AXT plans; for RCT plans created with
 * AXSPMV_PLAN_TRANSPOSE the position of each entry of the plan of A^T. */
int axspmv_plan_vbar_positions(const axspmv_plan *p, uint32_t *h_pos, uint64_t count);
/* Forward + adjoint of the loss 0.5*||A x - b||^2: y = A x, r = y - b,
 * x_bar = A^T r, v_bar[k] = r[row(k)] * x[col(k)].  fused = 1: one sweep over
 * A produces y, r and v_bar, then one sweep over A^T; fused = 0: SpMV,
 * residual, adjoint and v_bar as four separate kernels.  Needs
 * AXSPMV_PLAN_TRANSPOSE; plans: CSR naive, AXT UNC th == 1 and RCT.  RCT:
 * r needs every column slice's sums, so the forward SpMV and the residual
 * come first; fused = 1 then produces x_bar and v_bar in one sweep over the
 * plan of A^T (whose tiles keep their x values in LDS beside the x_bar sums;
 * with automatic tiling that plan uses 8192-row tiles and half the slices of
 * the forward rule); v_bar in the order of axspmv_vbar, written with 16-byte
 * stores where d_vbar is 16-byte aligned (8-byte aligned is enough). */
int axspmv_fwd_adj(axspmv_plan *p, const double *d_x, const double *d_b, double *d_y, double *d_r,
                   double *d_xbar, double *d_vbar, int fused, void *stream);

/* ---------------------------------------------------------------- SpMM */
/* Y = A X for k right-hand sides, X (ncols x k) and Y (nrows x k) row-major
 * ("interleaved": X[c*k + j]).  Every value/index load feeds ktile FMAs.
 * Plans: CSR, AXC, AXT UNC th == 1 and RCT (RCT: a tile keeps its rows x
 * ktile sums in LDS, so tile_rows * ktile * 8 bytes must fit one workgroup's
 * LDS -- tile_rows <= 2560 at ktile 8 on 160 KB; the default ktile is cut
 * to fit; with several column slices Y is cleared, then atomically added
 * to; 1024-lane workgroups unless the plan sets block).  No reference
 * counterpart. */
int axspmv_spmm(axspmv_plan *p, uint32_t k, const double *d_X, double *d_Y, void *stream);

/* ---------------------------------------------------------- timing */
typedef struct axspmv_timing {
    double mean_ms, median_ms, min_ms, max_ms; /* per operation */
    uint32_t iters, mode;                      /* the AXSPMV_TIME_* used */
    double main_ms; /* AXSPMV_TIME_KERNEL: mean time of the path's main
                       kernel alone; other modes: = mean_ms */
} axspmv_timing;
#define AXSPMV_TIME_EVENTS 0u /* one hipEvent pair + sync per launch (reference) */
#define AXSPMV_TIME_GRAPH 1u  /* replays of a hipGraph of `graph_batch` SpMVs    */
#define AXSPMV_TIME_REGION 2u /* one event pair around `iters` back-to-back SpMVs */
#define AXSPMV_TIME_COLD 4u   /* per-launch event pairs as EVENTS, each launch
                                 preceded by a 512 MB read that evicts the
                                 Infinity Cache and L2: times from HBM */
#define AXSPMV_TIME_KERNEL 3u /* as REGION (SpMV only, iters <= 4096), with an
                                 event before every SpMV and one between its
                                 launches: main_ms is the mean time of its
                                 main kernel.  Only RCT and AXT plans with
                                 column slices launch a second kernel: the
                                 merge of slice partials after the main one
                                 or the clear of y before it; other paths
                                 report main_ms = the SpMV's time.
                                 median/min/max: of the intervals between the
                                 starts of consecutive SpMVs in the region. */
/* Time `iters` SpMVs after `warmup` untimed ones.  AXSPMV_TIME_EVENTS is the
 * reference's loop (cudaSpmv.cu:563-581): mean and median of per-launch
 * times.  AXSPMV_TIME_GRAPH reports per-SpMV times of graph replays (for
 * kernels shorter than the launch latency).  AXSPMV_TIME_REGION reports the
 * average over one timed region of back-to-back launches (min = max = mean). */
int axspmv_time_spmv(axspmv_plan *p, const double *d_x, double *d_y, uint32_t iters,
                     uint32_t warmup, uint32_t mode, uint32_t graph_batch, void *stream,
                     axspmv_timing *out);

/* Operations axspmv_time_op can time, and their operands. */
enum axspmv_op {
    AXSPMV_OP_SPMV = 0,          /* y = A x                          */
    AXSPMV_OP_ADJOINT = 1,       /* xbar = A^T ybar                  */
    AXSPMV_OP_VBAR = 2,          /* vbar = ybar[row] * x[col]        */
    AXSPMV_OP_FWD_ADJ = 3,       /* axspmv_fwd_adj, fused            */
    AXSPMV_OP_FWD_ADJ_SPLIT = 4, /* axspmv_fwd_adj, four kernels     */
    AXSPMV_OP_SPMM = 5,          /* Y = A X, k right-hand sides      */
    AXSPMV_OP_XREFRESH = 6       /* AX-layout plan: only the x-half refresh
                                    of its ax array from x (the overhead
                                    time of the layout)               */
};
typedef struct axspmv_operands {
    const double *x, *b, *ybar, *X;
    double *y, *r, *xbar, *vbar, *Y;
    uint32_t k, pad_;
} axspmv_operands;
int axspmv_time_op(axspmv_plan *p, uint32_t op, const axspmv_operands *ops, uint32_t iters,
                   uint32_t warmup, uint32_t mode, uint32_t graph_batch, void *stream,
                   axspmv_timing *out);

/* ------------------------------------------------- builders on the device */
/* A CSR matrix whose arrays are in device memory (caller-owned): the same
 * rules as a host CSR (see axspmv_csr_check); the device builders check the
 * offsets and column bounds on the device. */
typedef struct axspmv_dcsr {
    uint32_t nrows, ncols, nnz, pad_;
    const uint32_t *row;  /* nrows + 1 */
    const uint32_t *col;  /* nnz */
    const double *val;    /* nnz */
} axspmv_dcsr;

/* A format built on the device.  Arrays are device memory owned by the
 * library (axspmv_dfmt_free); lengths and parameters as in axspmv_axc /
 * axspmv_axt.  AXC: con = brp (nrows + 1), plane = mapx (lenMAPX = lenAX/2 +
 * 8), sec unused.  AXT: con (lenCON + 1), sec (lenSEC), plane = column plane
 * (lenAX / 2).  ms = device time of the conversion (count, scan and fill
 * kernels; allocation excluded), the `ct` of the reference's formats table
 * (cudaSpmv.cu:2420-2436) measured on the GPU. */
typedef struct axspmv_dfmt {
    char name[48];
    uint32_t format, mode, hw, th, bs, log;
    uint32_t nrows, nnz, tileN, lenAX, lenSEC, lenCON, lenMAPX, pad_;
    double *ax;       /* lenAX: values, and x values when x was given (else 0) */
    uint32_t *sec;
    uint32_t *con;
    uint32_t *plane;
    double ms;
} axspmv_dfmt;

/* Device counterparts of axspmv_build_axt / axspmv_build_axc (the
 * reference's getFormatDataAXT / getFormatDataAXC, cudaSpmv.cu:1391-1566 and
 * 1004-1062, run on the host 250 times in the reference): the same arrays,
 * bit for bit, built by kernels from a device CSR.  d_x may be NULL (the
 * x-halves stay 0; axspmv_x_refresh fills them later).  Same parameter
 * rules and errors as the host builders.  The device CSR's offsets (row[0]
 * = 0, non-decreasing, row[nrows] = nnz) and column bounds are checked on
 * the device before anything is filled (AXSPMV_E_INVALID, nothing built);
 * any column order within a row is built as the host builders build it. */
int axspmv_dev_build_axt(const axspmv_dcsr *a, const double *d_x, uint32_t bs, uint32_t thw, uint32_t th,
                         uint32_t mode, void *stream, axspmv_dfmt **out);
int axspmv_dev_build_axc(const axspmv_dcsr *a, const double *d_x, uint32_t hbs, void *stream,
                         axspmv_dfmt **out);
void axspmv_dfmt_free(axspmv_dfmt *f);

/* axspmv_dev_build_axt without any device allocation.  No reference
 * counterpart: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  The caller passes the
 * output arrays in *f (ax, sec, con, plane; device memory it owns) with the
 * layout's lengths in f->lenAX, lenSEC, lenCON (= nrows) and lenMAPX (=
 * lenAX / 2) -- as an earlier axspmv_dev_build_axt of a matrix with the same
 * row lengths reported them (or axspmv_build_axt's lenAX / lenSEC); other
 * lengths are AXSPMV_E_INVALID, before anything is written.  The call fills
 * the arrays (bit for bit those of axspmv_dev_build_axt) and the other fields
 * of *f.  Its scratch (check flag, row counts, their scan, the scan's
 * temporary storage) lives in a static device array of
 * AXSPMV_DEV_BUILD_NOALLOC_SCRATCH bytes per GPU, held until the process
 * ends: UNC layouts need about 16 * (nrows + 1) bytes of it plus the scan's
 * storage (some 3 million rows at most), COM layouts only the flag; a matrix
 * that needs more is AXSPMV_E_UNSUPPORTED.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (threads or streams) race.  The call
 * synchronizes `stream` (it reads the layout's length back). */
#define AXSPMV_DEV_BUILD_NOALLOC_SCRATCH (48u << 20)
int axspmv_dev_build_axt_noalloc(const axspmv_dcsr *a, const double *d_x, uint32_t bs, uint32_t thw, uint32_t th,
                                 uint32_t mode, void *stream, axspmv_dfmt *f);

/* ------------------------------------------------------------ batches */
/* Many matrices of different sizes, SpMV in one launch.  x and y are the
 * concatenations of the matrices' x (ncols each) and y (nrows each) in batch
 * order.  opts[0].format AXSPMV_FMT_AXT (mode UNC): one persistent launch,
 * each matrix in the AXT UNC layout of its own options (opts[i] if
 * per_matrix, else opts[0] for all).  AXSPMV_FMT_RCT: the block-diagonal
 * matrix of all members as one RCT plan (one option set: per_matrix options
 * must all be equal; nitems in the info = its tiles); sizes of the whole
 * batch must fit 32 bits.  A graph of an RCT batch captures the plan's
 * kernels.  No reference counterpart. */
typedef struct axspmv_batch axspmv_batch;
int axspmv_batch_create(const axspmv_csr *const *mats, uint32_t count, const axspmv_opts *opts,
                        int per_matrix, void *stream, axspmv_batch **out);
int axspmv_batch_destroy(axspmv_batch *b);
typedef struct axspmv_batch_info {
    uint64_t nrows, ncols, nnz, slots, alg_bytes, device_bytes;
    uint32_t count, nitems, nwaves, pad_;
} axspmv_batch_info;
int axspmv_batch_get_info(const axspmv_batch *b, axspmv_batch_info *info);
int axspmv_batch_spmv(axspmv_batch *b, const double *d_x, double *d_y, void *stream);
/* Same modes as axspmv_time_op (AXSPMV_TIME_GRAPH captures graph_batch
 * persistent launches in one hipGraph). */
int axspmv_batch_time(axspmv_batch *b, const double *d_x, double *d_y, uint32_t iters, uint32_t warmup,
                      uint32_t mode, uint32_t graph_batch, void *stream, axspmv_timing *out);

/* ------------------------------------------------------------- graphs */
/* A hipGraph holding `repeats` back-to-back launches of one operation of a
 * plan (enum axspmv_op, operands as for axspmv_time_op) or of a batch SpMV,
 * captured once and replayed with axspmv_graph_launch.  The captured
 * kernels write every output element, so replays need no clearing.  No
 * reference counterpart. */
typedef struct axspmv_graph axspmv_graph;
int axspmv_plan_graph(axspmv_plan *p, uint32_t op, const axspmv_operands *ops, uint32_t repeats,
                      axspmv_graph **out);
int axspmv_batch_graph(axspmv_batch *b, const double *d_x, double *d_y, uint32_t repeats,
                       axspmv_graph **out);
int axspmv_graph_launch(axspmv_graph *g, void *stream);
int axspmv_graph_destroy(axspmv_graph *g);

/* ---------------------------------------------------------- autotuner */
/* Times every candidate option set (a plan built on the device, then `iters`
 * back-to-back launches of `op`, AXSPMV_OP_SPMV or AXSPMV_OP_SPMM with k
 * right-hand sides) and returns the index of the fastest in *best.
 * times_ms (ncand entries, may be NULL) gets each candidate's time per
 * operation, or -1 for a candidate the library rejects.  cand = NULL uses
 * the built-in set of axspmv_autotune_defaults; ncand then gives the room in
 * times_ms (which must hold the whole set).  No reference counterpart: the
 * reference sweeps block sizes by relaunching (jobLauncher.sh:145-176). */
/* The built-in candidates: for SpMV AXT UNC th = 1 with thw 2..16 and th =
 * 2..8 with thw 64 at two block budgets, AXT UNC th = 1 (thw 2) in 2, 4 and
 * 8 column slices, K1 (chunk 64), RCT (default, 256 and 1024 threads), and
 * RCT and AXT UNC th = 1 (thw 4) with AXSPMV_PLAN_REORDER (rejected, time
 * -1, for non-square matrices); for batches the AXT part and the default RCT options;
 * for SpMM the four AXT k-tiles and eleven RCT plans
 * (ktile 8 with 1024 and 2048 rows, 16 with 512, 32 with 256, each with one
 * column slice and with the default slicing; and ktile 8 / 16 / 32 with 512
 * / 256 / 128 rows and one slice).  Writes min(cap, count) options to out
 * and the set's size to *count. */
int axspmv_autotune_defaults(uint32_t op, axspmv_opts *out, uint32_t cap, uint32_t *count);
int axspmv_autotune(const axspmv_csr *a, uint32_t op, uint32_t k, const axspmv_opts *cand,
                    uint32_t ncand, uint32_t iters, void *stream, uint32_t *best, double *times_ms);
/* Options for a batch.  AXT UNC candidates are chosen per matrix: matrices
 * are grouped by shape (rows rounded to a power of two, mean row length in
 * steps of 4) and each group is tuned on a batch of its own members
 * (repeated until it holds about 4M nonzeros, so a candidate is timed at the
 * throughput it reaches inside a large batch).  RCT candidates apply to the
 * whole batch (one block-diagonal plan): the whole batch is timed with the
 * per-group AXT choice and with each RCT option set, and the fastest is
 * written to out_opts (every entry the same option set if RCT wins).
 * *ngroups gets the number of distinct option sets chosen.  cand = NULL uses
 * the AXT UNC part of the built-in SpMV set plus the default RCT options. */
int axspmv_autotune_batch(const axspmv_csr *const *mats, uint32_t count, const axspmv_opts *cand,
                          uint32_t ncand, uint32_t iters, void *stream, axspmv_opts *out_opts,
                          uint32_t *ngroups);

/* ---------------------------------------------------- device helpers */
int axspmv_dev_alloc(size_t bytes, void **ptr);
int axspmv_dev_free(void *ptr);
int axspmv_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int axspmv_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int axspmv_stream_sync(void *stream);
int axspmv_device_sync(void);
/* Number of visible HIP devices (0 if none). */
int axspmv_device_count(int *count);
/* Plain device copy dst[i] = src[i] (a bandwidth reference kernel). */
int axspmv_copy_f64(const double *d_src, double *d_dst, size_t n, void *stream);

/* The x-refresh step of the interleaved layouts: for q < nslots,
 *   d_ax[(q / hw) * 2hw + hw + q % hw] = d_map[q] < nx ? d_x[d_map[q]] : 0.0
 * i.e. fills the x-half of every (hw values, hw x-values) group of an AXC
 * (hw = hbs) or AXT (hw = thw; th > 1 levels and COM tiles are such groups
 * too) ax array from the column of each value slot.  d_map is the AXC mapx
 * (get_mapxAXC, cudaSpmv.cu:983-1000) or an AXT column plane; their padding
 * slots hold column 0 and so get x[0], as the reference's own gather of the
 * x-halves does (the AVX-512 x_ax, gpuSpmv_header.h:1750-1779), while a map
 * entry >= nx (e.g. 0xffffffff) marks padding that gets 0.0, which makes
 * the result equal the builders' ax.  hw is a power of two <= 1024 and
 * nslots a multiple of hw (d_ax holds 2 * nslots doubles). */
int axspmv_x_refresh(const uint32_t *d_map, uint64_t nslots, uint32_t hw, uint32_t nx, const double *d_x,
                     double *d_ax, void *stream);

/* The wave64 cross-lane primitives that replace the reference's 32-lane
 * __shfl_* reduction trees (cudaSpmv.cu:1086-1090, 1586-1594) and its
 * Hillis-Steele scans (cudaSpmv.cu:2061-2086, scan.cu:239-414), run lane by
 * lane so that a caller can check them: element i is lane i % 64 of a wave,
 * n a multiple of 64, and for each op
 *   SEG_SCAN     ov = inclusive sum of v from the lane's segment start (the
 *                last lane <= i of its wave with h != 0, else lane 0);
 *                oh != 0 iff some lane <= i of its wave has h != 0
 *   SCAN_MAX     oh = max of h over lanes 0..i of the wave; ov = v
 *   SHIFT_UP     ov, oh = v, h of lane i - 1; lane 0 gets -1.0, 0xffffffff
 *   SHIFT_DOWN   oh = h of lane i + 1; lane 63 gets 0xffffffff; ov = v
 *   GROUP_SUM_G  ov = sum of v over the aligned group of G lanes holding
 *                lane i (G = 1 << (op - AXSPMV_WAVE_GROUP_SUM_1)); oh = h
 * Test hook only: no SpMV path calls it. */
enum axspmv_wave_op {
    AXSPMV_WAVE_SEG_SCAN = 0,
    AXSPMV_WAVE_SCAN_MAX = 1,
    AXSPMV_WAVE_SHIFT_UP = 2,
    AXSPMV_WAVE_SHIFT_DOWN = 3,
    AXSPMV_WAVE_GROUP_SUM_1 = 4,
    AXSPMV_WAVE_GROUP_SUM_64 = 10
};
int axspmv_wave_selftest(uint32_t op, const double *d_v, const uint32_t *d_h, double *d_ov, uint32_t *d_oh,
                         uint64_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* AXSPMV_H */
