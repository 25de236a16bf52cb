// Host-side launch interface of the HIP kernels (implemented in the .hip
// files).  Only plain pointers and sizes cross this boundary.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace axspmv {
namespace launch {

// A wave block: units [beg, end) of one wave (row-aligned for UNC / CSR).
struct Block {
    uint32_t beg, end;
};

// Columns a wave block touches, [cmin, cmin + width): with width <=
// STAGE_COLS the wave stages that part of x in LDS before its gathers.
struct Window {
    uint32_t cmin, width;
};
constexpr uint32_t STAGE_COLS = 1024;

// Row-aligned wave-block kernels (AXT UNC th==1 / th>1, CSR stream).
struct SegArgs {
    const void *vplane;       // values (double or float by vtype)
    const uint32_t *cplane;   // column of every slot
    const uint32_t *rowsrc;   // sec[] (UNC) or row pointers (CSR stream)
    const Block *blocks;      // nblocks wave blocks
    uint32_t nblocks;
    const uint32_t *empty;    // rows without entries (y = 0)
    uint32_t nempty;
    uint32_t thw, th;         // AXT tile half width / height
    uint32_t vtype;           // 0 = f64 values, 1 = f32 values
    uint32_t block;           // threads per workgroup
    const Window *win;        // per block (nullptr: x from HBM for every block)
    uint32_t nrows_add;       // AXT UNC th == 1 in column slices: y (this many
                              // rows) is cleared, then every block adds its
                              // rows' sums; 0: blocks store whole rows
    hipEvent_t mid;           // column slices with atomics: recorded between
                              // the clear and the block kernel (nullptr: none)
    double *part;             // column slices, deterministic: S x nrows_add
                              // partial sums (empty lists s * nrows + r of the
                              // rows without entries in slice s), summed in
                              // slice order by a merge kernel
    uint32_t S;               // column slices (with part)
    uint32_t nt = 0;          // AXT UNC: 1 streams the planes with
                              // non-temporal loads
};
hipError_t axt_unc_h1(const SegArgs &a, const double *x, double *y, hipStream_t s);
// As axt_unc_h1 on the interleaved layout: a.vplane is the FP64 ax array
// (tiles of thw values then thw x-values, x-halves already refreshed).
hipError_t axt_unc_h1_ax(const SegArgs &a, double *y, hipStream_t s);
// ax[(q >> hw_log) * 2hw + hw + q % hw] = map[q] < nx ? x[map[q]] : 0 for q < n.
hipError_t x_refresh(const uint32_t *map, uint64_t n, uint32_t hw_log, uint32_t nx, const double *x, double *ax,
                     hipStream_t s);
hipError_t axt_unc_h(const SegArgs &a, const double *x, double *y, hipStream_t s);

// Plane slot -> CSR position of the AXT UNC th == 1 layout.
struct VbarArgs {
    const uint32_t *rowptr;  // CSR row pointer, nrows + 1
    const uint32_t *first;   // first plane slot of each row
};
hipError_t axt_unc_h1_vbar(const SegArgs &a, uint32_t nunits, const double *x, const double *ybar, double *vbar,
                           const VbarArgs &va, hipStream_t s);
// a.blocks must hold blocks of at most 64 units or single rows.
hipError_t axt_unc_h1_fused(const SegArgs &a, const double *x, const double *b, double *y, double *res,
                            double *vbar, const VbarArgs &va, hipStream_t s);
hipError_t csr_vbar(uint32_t nrows, const uint32_t *col, const uint32_t *row, const double *x, const double *ybar,
                    double *vbar, hipStream_t s);
hipError_t csr_fused(uint32_t nrows, const void *val, uint32_t vtype, const uint32_t *col, const uint32_t *row,
                     const double *x, const double *b, double *y, double *res, double *vbar, hipStream_t s);
hipError_t residual(uint32_t n, const double *y, const double *b, double *r, hipStream_t s);

// SpMM on a row-contiguous layout: row r = slots first[r] .. first[r] + rl - 1.
struct SpmmArgs {
    uint32_t nrows, vtype, long_slots;
    const uint32_t *first, *rowptr;
    const void *val;
    const uint32_t *col;
    const uint32_t *longrows;
    uint32_t nlong;
};
// kt (k-tile) in {8, 16, 32, 64}
hipError_t spmm(const SpmmArgs &a, uint32_t k, uint32_t kt, const double *X, double *Y, hipStream_t s);

// RCT tiles: row block rb = rows rb_start[rb] .. rb_start[rb+1]-1 (at most R
// rows); tile t = rb * S + s holds entries tile_ptr[t] .. tile_ptr[t+1]-1
// (column col, row rb_start[rb] + lrow), sorted by column.
struct RctArgs {
    const void *val;
    const uint32_t *col;    // unpacked: column of each entry
    const uint16_t *lrow;   // unpacked: row within the tile
    const uint32_t *key;    // packed: offset << rbits | row, where entry e's
                            // column is (sbase[e >> 10] + offset) mod ncols
    const uint32_t *sbase;  // packed: base column of every 1024-entry segment
    const uint32_t *tile_ptr;
    const uint32_t *rb_start;  // nrb + 1 row-block starts
    uint32_t nrows, R, S, nrb, vtype, block, rbits, ncols;
    double *part;  // S > 1: S x nrows slice partials (nullptr: atomic adds into y)
    uint32_t nt_stream = 0;  // 1: the tile kernel (SpMV, value-gradient sweep) streams with non-temporal loads
    // narrow bands: a 16-bit key per entry (key16 != nullptr; key, col and
    // lrow unused) with the column as an offset of cbits16 bits from sbase
    // (runs of 128 entries) and row - column - dmin above it (dmin modulo
    // 2^32)
    const uint16_t *key16 = nullptr;
    uint32_t cbits16 = 0, dmin = 0;
};
// S > 1 with a.part: the tile kernel stores every slice's row sums in its
// partial vector, then part_merge adds them in slice order into y; mid
// (optional) is recorded on s between the two.  S > 1 without a.part: a fill
// kernel clears y, then the tile kernel adds into it; mid is recorded
// between those.  vbar != nullptr (a plan of A^T, 16 B of LDS per tile row,
// no partials): also vbar[e] = xr[row(e)] * x[col(e)] for every entry, in
// entry order.  res != nullptr (vbar == nullptr): also res = y - rb, written
// where y's rows are (the merge, or one slice's tile kernel; after the adds,
// a pass of its own).
hipError_t rct(const RctArgs &a, const double *x, double *y, hipStream_t s, hipEvent_t mid = nullptr,
               const double *xr = nullptr, double *vbar = nullptr, const double *rb = nullptr,
               double *res = nullptr);
// Y = A X (k interleaved right-hand sides) on a one-slice plan, kt (8, 16,
// 32 or 64) right-hand sides per launch; needs R * kt * 8 bytes of LDS
hipError_t rct_spmm(const RctArgs &a, uint32_t k, uint32_t kt, const double *X, double *Y, hipStream_t s);
// vbar[e] = ybar[row(e)] * x[col(e)] for every entry e, in entry (tile) order
hipError_t rct_vbar(const RctArgs &a, const double *ybar, const double *x, double *vbar, hipStream_t s);

// Batched persistent SpMV over AXT UNC matrices.
struct BatchMat {
    uint64_t voff;      // first slot of the matrix in the concatenated planes
    uint64_t soff;      // first entry of its sec[] in the concatenated sec
    uint32_t xoff, yoff;  // offsets of its x / y in the concatenated vectors
    uint32_t kind;      // 0/1/2: th == 1 with 4/2/1 slots per unit; 3: th > 1
    uint32_t ushift;    // th == 1: log2(thw / slots per unit)
    uint32_t thw_log, th;
};
struct BatchItem {
    uint32_t mat, beg, end, pad;
};
struct BatchArgs {
    const void *vplane;
    const uint32_t *cplane, *sec;
    const BatchMat *mats;
    const BatchItem *items;
    uint32_t nitems, nwaves, vtype;
    const uint32_t *empty;  // global y indices of rows without entries
    uint32_t nempty;
    const Window *win;      // per item (nullptr: x from HBM for every item)
};
hipError_t batch(const BatchArgs &a, const double *x, double *y, hipStream_t s);

hipError_t csr_naive(uint32_t nrows, const void *val, uint32_t vtype, const uint32_t *col, const uint32_t *row,
                     const double *x, double *y, uint32_t block, hipStream_t s);
// lanes (1..64, a power of two) per row; rows of more than long_len entries
// are listed in longrows and done one wave per row
hipError_t csr_vector(uint32_t nrows, const void *val, uint32_t vtype, const uint32_t *col, const uint32_t *row,
                      uint32_t lanes, uint32_t long_len, const uint32_t *longrows, uint32_t nlong, const double *x,
                      double *y, hipStream_t s, uint32_t nt = 0);

// K1: a work item is steps [j0, j1) of one chunk; slot == 0xffffffff writes
// y directly, otherwise partial sums go to part[slot*chunk + r] and the
// split list (chunk, first slot, slot count) sums them.
struct K1Item {
    uint32_t chunk, chp, j0, j1, slot;
};
struct K1Args {
    const K1Item *items;
    uint32_t nitems;
    const uint32_t *split;  // 3 words per split chunk
    uint32_t nsplit;
    uint32_t chunk, nrows, vtype;
    const void *val;
    const uint32_t *col, *permi;
    double *part;
    uint32_t nt = 0;  // 1: val and col streamed with non-temporal loads
};
hipError_t k1(const K1Args &a, const double *x, double *y, hipStream_t s);

// AXC: groups of hbs lanes per row; rows longer than long_slots slots are
// listed in longrows and done one wave per row.
struct AxcArgs {
    uint32_t nrows, hbs, vtype, long_slots;
    const uint32_t *brp, *mapx;
    const void *vplane;
    const uint32_t *longrows;
    uint32_t nlong;
    uint32_t nt = 0;  // 1: vplane and mapx streamed with non-temporal loads
};
hipError_t axc(const AxcArgs &a, const double *x, double *y, hipStream_t s);

// AXT COM: a wave owns slots (th == 1) or tile columns (th > 1) [beg, end);
// flags bit 0: its first run continues a row from the previous wave, bit 1:
// its last run continues into the next wave.
struct ComWave {
    uint32_t beg, end, flags, pad;
};
struct ComArgs {
    const void *vplane;
    const uint32_t *cplane, *hdr;
    const ComWave *waves;
    uint32_t nwaves, thw, th, bs, log, nnz, vtype;
    double *head, *tail;        // per-wave partials of spanning rows
    const uint32_t *span;       // (row, first wave, last wave) per spanning row
    uint32_t nspan;
    const uint32_t *empty;      // rows without entries
    uint32_t nempty;
    uint32_t nt = 0;            // 1: planes streamed with non-temporal loads
};
hipError_t axt_com(const ComArgs &a, const double *x, double *y, hipStream_t s);

// Format builders on the device (CSR arrays already in device memory).
struct DevCsr {
    uint32_t nrows, ncols, nnz;
    const uint32_t *row, *col;
    const double *val;
};
struct BuildGeo {
    uint32_t hw, th, bs, log;  // tile half width (AXC: hbs), tile height, COM block size, header shift
};
struct BuildOut {
    double *ax;        // interleaved values / x values (zero filled beforehand)
    uint32_t *plane;   // column of every value slot (AXC: mapx)
    uint32_t *sec;     // AXT row / header words (unused by AXC)
    uint32_t *con;     // AXT con, AXC brp
};
// cnt[r] = ceil(rl_r / div) for r < nrows
// the offsets / column checks of check_kernel OR-ed into *flag (bits 1, 2,
// 4, 8; the caller clears it first)
hipError_t build_check(const DevCsr &a, uint32_t *flag, hipStream_t s);
hipError_t build_count(const DevCsr &a, uint32_t div, uint64_t *cnt, hipStream_t s);
// rocPRIM exclusive scan; tmp == nullptr queries tmp_bytes
hipError_t build_scan(const uint64_t *in, uint64_t *out, size_t n, void *tmp, size_t &tmp_bytes, hipStream_t s);
// kind 0: AXC / AXT UNC th == 1, 1: AXT UNC th > 1, 2: COM th == 1, 3: COM th > 1;
// start = scanned counts (nrows + 1 entries) for kinds 0 and 1
hipError_t build_fill(const DevCsr &a, int kind, const uint64_t *start, const BuildGeo &g, const double *x,
                      const BuildOut &o, hipStream_t s);

hipError_t copy_f64(const double *src, double *dst, size_t n, hipStream_t s);
hipError_t fill_f64(double *dst, size_t n, double v, hipStream_t s);
hipError_t fill_u32(uint32_t *dst, size_t n, uint32_t v, hipStream_t s);
// y[0 .. n) = 0 in one pass (16-byte stores where y is 16-byte aligned)
hipError_t clear_f64(double *y, size_t n, hipStream_t s);
// y[r] = part[r] + part[nrows + r] + ... + part[(S - 1) nrows + r], in that
// order; with res, also res[r] = y[r] - rb[r]
hipError_t part_merge(const double *part, uint32_t S, uint32_t nrows, double *y, hipStream_t s,
                      const double *rb = nullptr, double *res = nullptr);
// reads p[0 .. n) (n even, p 16-byte aligned): evicts caches without dirtying them
hipError_t touch(const double *p, size_t n, double *sink, hipStream_t s);
// the wave64 primitive op (enum axspmv_wave_op) lane by lane; n a multiple of 64
hipError_t wave_selftest(uint32_t op, const double *v, const uint32_t *h, double *ov, uint32_t *oh, size_t n,
                         hipStream_t s);
// For every chunk c of L positions of src (n in all; the last chunk may be
// short; L <= 65536 and L * 8 bytes of LDS): lds[i] = src[c L + i], then
// dst[w ? w[c L + i] : c L + i] = lds[b[c L + i]] (every w value < the length
// of dst, every b value < the chunk's length)
hipError_t permute_chunks(const uint32_t *w, const uint16_t *b, uint32_t n, uint32_t L, const double *src,
                          double *dst, hipStream_t s);
// For every chunk c of L positions (n in all; L <= 10240: 2 L * 8 bytes of
// LDS): y[c L + i] = t[c L + b2[c L + i]], r[c L + i] = y[c L + i] -
// bv[c L + i], then t2[w1[c L + i]] = r[c L + b1[c L + i]] -- the second pass
// of one permutation and the first of another through the same chunks,
// joined by a residual (t2 must not overlap t)
hipError_t back_residual_in(const uint16_t *b2, const uint32_t *w1, const uint16_t *b1, uint32_t n, uint32_t L,
                            const double *t, const double *bv, double *y, double *r, double *t2, hipStream_t s);
// dst[i k + j] = src[idx[i] k + j] for i < n, j < k (every idx[i] k + k <= the
// length of src)
hipError_t gather_rows(const uint32_t *idx, uint32_t n, uint32_t k, const double *src, double *dst, hipStream_t s);
// y[rows[i]] = 0 for i < n
hipError_t fill_rows(const uint32_t *rows, uint32_t n, double *y, hipStream_t s);

}  // namespace launch
}  // namespace axspmv
