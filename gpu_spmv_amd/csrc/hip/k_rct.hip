// RCT: row-block x column-slice tiles with entries sorted by column.
//
// Random column indices make SpMV a gather problem: every nonzero costs one
// L2 request for x, and once x outgrows an XCD's 4 MB L2 the requests go to
// the Infinity Cache.  RCT attacks both:
//
//  * the matrix is cut into tiles of (at most) R rows x C columns (S = ncols /
//    C column slices); row blocks hold about equal numbers of nonzeros (the
//    host cuts them), and a workgroup owns one tile and accumulates its rows
//    in LDS;
//  * entries of a tile are stored sorted by column, so neighbouring lanes
//    gather neighbouring x entries -- one cache line serves several lanes;
//  * workgroups are dispatched round-robin to the 8 XCDs (workgroup w runs on
//    XCD w % 8), and tile w belongs to column slice (w % 8) % S: each XCD only
//    ever reads its own slice of x, which stays resident in its L2.
//
// With one slice a tile's sums are its rows' totals and are stored to y.
// With S slices either every tile stores its rows' sums in its slice's
// partial vector (S x n doubles held by the plan) and a merge kernel adds the
// S partials of each row in slice order, or y is cleared first and every
// tile adds its sums into it with FP64 atomics (global_atomic_add_f64).
// Every tile writes its rows when its loop ends and all tiles end together;
// the host picks partials where that burst is heavy against the loop.  The
// LDS adds (ds_add_f64) make the order of the additions inside a row
// nondeterministic; the result stays within the FP64 tolerance of the
// left-to-right CPU sum.
#include "launch.hpp"
#include "wave64.hpp"

// Entries per lane per round (each of the two register sets holds this many,
// as RCT_UNROLL / 2 pairs in the tile kernel).
#ifndef RCT_UNROLL
#define RCT_UNROLL 4
#endif

namespace {
using axspmv::launch::RctArgs;

// Entry e of a tile: value val[e]; its column and tile row stored one of
// three ways (KEY):
//  * KEY_SPLIT: col[e] and lrow[e];
//  * KEY_PACKED: one word key[e] = offset << rbits | row, the column being
//    (sbase[e >> 10] + offset) mod ncols: every run of 1024 entries has a
//    base column (its first entry's or its smallest), and its entries,
//    sorted by column tile by tile, lie a short way past it -- going round
//    from the last column to column 0 where a run passes from one row block
//    to the next;
//  * KEY_DIAG (narrow bands): a 16-bit key16[e] = d << cbits | offset, the
//    column being sbase[e >> 7] + offset (the smallest column of its run of
//    128 entries) and the row (global) column + d + dmin, where d + dmin is
//    the entry's row minus its column; the tile row is that row minus the
//    tile's first row (rofs = dmin - row0, all modulo 2^32).
// NT: the paired stream loads of load2 are non-temporal (the plan sets it
// when the matrix stream is too large to stay in the Infinity Cache from one
// SpMV to the next; there the L2 lines it would leave behind only crowd out
// x).
enum : int { KEY_SPLIT = 0, KEY_PACKED = 1, KEY_DIAG = 2 };

template <class VT, int KEY, bool NT = false>
struct RctEntries {
    const VT *__restrict__ val;
    const uint32_t *__restrict__ key;
    const uint32_t *__restrict__ col;
    const uint16_t *__restrict__ lrow;
    const uint16_t *__restrict__ key16;
    const uint32_t *__restrict__ sbase;
    uint32_t rbits, ncols, cbits, rofs;
    // the tile whose rows start at row0
    __device__ __forceinline__ RctEntries(const RctArgs &a, uint32_t row0)
        : val(static_cast<const VT *>(a.val)), key(a.key), col(a.col), lrow(a.lrow), key16(a.key16),
          sbase(a.sbase), rbits(a.rbits), ncols(a.ncols), cbits(a.cbits16), rofs(a.dmin - row0) {}
    __device__ __forceinline__ uint32_t column(uint32_t base, uint32_t off) const {
        const uint32_t room = ncols - base;  // columns from base to the end
        return off >= room ? off - room : base + off;
    }
    __device__ __forceinline__ void diag(uint32_t base, uint32_t k, uint32_t &c, uint32_t &r) const {
        c = base + (k & ((1u << cbits) - 1));
        r = c + (k >> cbits) + rofs;
    }
    __device__ __forceinline__ void load(uint32_t e, uint32_t &c, uint32_t &r, double &v) const {
        v = double(val[e]);
        at(e, c, r);
    }
    // entries 2q and 2q + 1 with one 16-byte (8 for FP32 values) and one
    // 8-byte (KEY_DIAG: 4-byte) load; the plan pads the arrays to an even
    // number of entries
    __device__ __forceinline__ void load2(uint32_t q, uint32_t (&c)[2], uint32_t (&r)[2], double (&v)[2]) const {
        typedef double f64x2 __attribute__((ext_vector_type(2)));
        typedef float f32x2 __attribute__((ext_vector_type(2)));
        typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
        if constexpr (sizeof(VT) == 8) {
            const f64x2 *p = reinterpret_cast<const f64x2 *>(val) + q;
            const f64x2 vv = NT ? __builtin_nontemporal_load(p) : *p;
            v[0] = vv.x;
            v[1] = vv.y;
        } else {
            const f32x2 *p = reinterpret_cast<const f32x2 *>(val) + q;
            const f32x2 vv = NT ? __builtin_nontemporal_load(p) : *p;
            v[0] = double(vv.x);
            v[1] = double(vv.y);
        }
        if constexpr (KEY == KEY_DIAG) {
            const uint32_t *pk = reinterpret_cast<const uint32_t *>(key16) + q;
            const uint32_t k = NT ? __builtin_nontemporal_load(pk) : *pk;
            const uint32_t b = sbase[q >> 6];  // entries 2q, 2q + 1: one run of 128
            diag(b, k & 0xffffu, c[0], r[0]);
            diag(b, k >> 16, c[1], r[1]);
        } else if constexpr (KEY == KEY_PACKED) {
            const u32x2 *pk = reinterpret_cast<const u32x2 *>(key) + q;
            const u32x2 k = NT ? __builtin_nontemporal_load(pk) : *pk;
            const uint32_t m = (1u << rbits) - 1, b = sbase[q >> 9];  // entries 2q, 2q + 1: one segment
            c[0] = column(b, k.x >> rbits);
            r[0] = k.x & m;
            c[1] = column(b, k.y >> rbits);
            r[1] = k.y & m;
        } else {
            const u32x2 *pc = reinterpret_cast<const u32x2 *>(col) + q;
            const uint32_t *pr = reinterpret_cast<const uint32_t *>(lrow) + q;
            const u32x2 cc = NT ? __builtin_nontemporal_load(pc) : *pc;
            const uint32_t rr = NT ? __builtin_nontemporal_load(pr) : *pr;
            c[0] = cc.x;
            c[1] = cc.y;
            r[0] = rr & 0xffffu;
            r[1] = rr >> 16;
        }
    }
    // column and tile row only
    __device__ __forceinline__ void at(uint32_t e, uint32_t &c, uint32_t &r) const {
        if constexpr (KEY == KEY_DIAG) {
            diag(sbase[e >> 7], key16[e], c, r);
        } else if constexpr (KEY == KEY_PACKED) {
            const uint32_t k = key[e];
            c = column(sbase[e >> 10], k >> rbits);
            r = k & ((1u << rbits) - 1);
        } else {
            c = col[e];
            r = lrow[e];
        }
    }
};

int key_mode(const RctArgs &a) { return a.key16 ? KEY_DIAG : a.key ? KEY_PACKED : KEY_SPLIT; }

// Tile w of the grid: XCD w % 8 takes slice (w % 8) % S; false for the
// workgroups past the last row block.
struct TileOf {
    uint32_t s, rb;
    __device__ __forceinline__ bool from(const RctArgs &a, uint32_t w) {
        const uint32_t xcd = w % 8, i = w / 8;
        const uint32_t per = 8 / a.S;  // row-block groups per slice
        s = xcd % a.S;
        rb = i * per + xcd / a.S;
        return rb < a.nrb;
    }
};

// VB: the same product on the plan of A^T, fused with the value gradient of
// A: vb[e] = xr[row(e)] * x[col(e)] for every entry, in entry order (the
// tile's xr values sit in LDS after its sums).  With x = r and xr = the
// forward x, y = A^T r and vb holds r[i] * x[j] for every nonzero A(i, j).
template <class VT, int KEY, bool VB, bool NT>
__global__ void __launch_bounds__(1024) rct_kernel(RctArgs a, const double *__restrict__ x, double *__restrict__ y,
                                                   const double *__restrict__ xr, double *__restrict__ vb,
                                                   const double *__restrict__ bvec, double *__restrict__ res) {
    extern __shared__ double yl[];
    TileOf tile;
    if (!tile.from(a, blockIdx.x)) return;  // the whole workgroup leaves together
    const uint32_t s = tile.s, rb = tile.rb;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t row0 = a.rb_start[rb], rows = a.rb_start[rb + 1] - row0;  // rows <= a.R
    double *xs = yl + a.R;  // VB: xr of the tile's rows
    for (uint32_t k = tid; k < rows; k += nt) {
        yl[k] = 0.0;
        if constexpr (VB) xs[k] = xr[row0 + k];
    }
    __syncthreads();
    const uint32_t t = rb * a.S + s;  // tile index in storage order
    const uint32_t beg = a.tile_ptr[t], end = a.tile_ptr[t + 1];
    const RctEntries<VT, KEY, NT> ent(a, row0);
    // Each lane takes entries in aligned pairs (2q, 2q + 1): a value load
    // moves 16 bytes and a key load 8, and the value gradient (VB) goes out
    // as one 16-byte store per pair.  Pairs pb .. pl cover the tile and each
    // holds at least one of its entries; an entry outside the tile borrows
    // its partner's column and row (both in the tile) and adds nothing.
    // Measured against one entry per lane: SpMV 1-4 % faster on the
    // power-law, stencil and banded matrices, forward + adjoint 2-4 %.
    //
    // Rounds of UP pairs per lane, in two register sets A and B: a round
    // issues its x gathers (and, for VB, its LDS reads of xr), then the loads
    // of the next round into the other set, then waits for the gathers only
    // and does its LDS adds.  Loads are never predicated (a lane past the end
    // re-reads the tile's last pair) and the products of such a lane are
    // zeroed.  The round count is the same for the whole workgroup, and the
    // last round (which fetches nothing) is peeled, so every fetch is used on
    // every path and the compiler has no reason to sink it behind the LDS
    // adds.
    //
    // VB stores are unpredicated too: a lane's two value gradients go to
    // entries of the tile (an entry outside it, or a lane past the end,
    // stores its partner's own value again), as one 16-byte store when every
    // lane of the wave holds a whole pair and vb is 16-byte aligned, else as
    // two 8-byte stores; non-temporal, as nothing here reads vb again (the
    // permuted 128^3 stencil's forward + adjoint, renumbered: 485.8 ->
    // 478.8 us; the 2^20-row power-law matrix's: 233.0 -> 229.2 us).
    constexpr int UP = (RCT_UNROLL + 1) / 2;
    struct PSet {
        uint32_t c[UP][2], r[UP][2];
        uint32_t at[UP][2];  // VB: the entries the two stores write
        double v[UP][2];
        bool ok[UP][2];      // adds its product
    };
    if (beg < end) {
        const uint32_t pb = beg >> 1, pl = (end - 1) >> 1, span = UP * nt;
        const uint32_t nr = (pl - pb + span) / span;
        const bool vb16 = (reinterpret_cast<uintptr_t>(vb) & 15) == 0;
        auto fetch = [&](PSet &S, uint32_t q0) {
#pragma unroll
            for (int j = 0; j < UP; ++j) {
                const uint32_t q = q0 + j * nt, qc = min(q, pl);
                ent.load2(qc, S.c[j], S.r[j], S.v[j]);
                const bool in0 = 2 * qc >= beg, in1 = 2 * qc + 1 < end;
                const uint32_t c0 = S.c[j][0], r0 = S.r[j][0], c1 = S.c[j][1], r1 = S.r[j][1];
                S.c[j][0] = in0 ? c0 : c1;
                S.r[j][0] = in0 ? r0 : r1;
                S.c[j][1] = in1 ? c1 : c0;
                S.r[j][1] = in1 ? r1 : r0;
                S.ok[j][0] = q == qc && in0;  // a lane past the end adds nothing
                S.ok[j][1] = q == qc && in1;
                if constexpr (VB) {
                    S.at[j][0] = in0 ? 2 * qc : 2 * qc + 1;
                    S.at[j][1] = in1 ? 2 * qc + 1 : 2 * qc;
                }
            }
        };
        auto gather = [&](const PSet &S, double (&xv)[UP][2], double (&xl)[UP][2]) {
#pragma unroll
            for (int j = 0; j < UP; ++j) {
                xv[j][0] = x[S.c[j][0]];
                xv[j][1] = x[S.c[j][1]];
            }
            if constexpr (VB) {
#pragma unroll
                for (int j = 0; j < UP; ++j) {
                    xl[j][0] = xs[S.r[j][0]];
                    xl[j][1] = xs[S.r[j][1]];
                }
            }
        };
        auto add = [&](const PSet &S, const double (&xv)[UP][2], const double (&xl)[UP][2]) {
#pragma unroll
            for (int j = 0; j < UP; ++j) {
#pragma unroll
                for (int h = 0; h < 2; ++h) atomicAdd(&yl[S.r[j][h]], S.ok[j][h] ? S.v[j][h] * xv[j][h] : 0.0);
                if constexpr (VB) {
                    const double p0 = xl[j][0] * xv[j][0], p1 = xl[j][1] * xv[j][1];
                    if (__all(vb16 && S.at[j][1] == S.at[j][0] + 1)) {  // whole pairs (at[j][0] even)
                        typedef double f64x2 __attribute__((ext_vector_type(2)));
                        __builtin_nontemporal_store(f64x2{p0, p1}, reinterpret_cast<f64x2 *>(vb) + (S.at[j][0] >> 1));
                    } else {
                        __builtin_nontemporal_store(p0, vb + S.at[j][0]);
                        __builtin_nontemporal_store(p1, vb + S.at[j][1]);
                    }
                }
            }
        };
        auto round = [&](const PSet &S, PSet &N, uint32_t q0) {
            double xv[UP][2], xl[UP][2];
            gather(S, xv, xl);
            fetch(N, q0 + span);
            add(S, xv, xl);
        };
        auto final_round = [&](const PSet &S) {
            double xv[UP][2], xl[UP][2];
            gather(S, xv, xl);
            add(S, xv, xl);
        };
        PSet A, B;
        uint32_t q = pb + tid;
        fetch(A, q);
        uint32_t k = 1;  // rounds k - 1 and k run in one pass; both have a successor
        for (; k + 1 < nr; k += 2) {
            round(A, B, q);
            round(B, A, q + span);
            q += 2 * span;
        }
        if (k < nr) {
            round(A, B, q);
            final_round(B);
        } else {
            final_round(A);
        }
    }
    __syncthreads();
    double *out = y + row0;
    if (a.S == 1) {
        // non-temporal: y kept in the L2s would take room from the next
        // SpMV's x (the 1024^2 Laplacian: 14.6 -> 13.9 us)
        for (uint32_t k = tid; k < rows; k += nt) {
            __builtin_nontemporal_store(yl[k], out + k);
            if (res) res[row0 + k] = yl[k] - bvec[row0 + k];  // forward + adjoint: r = y - b, read next
        }
    } else if (a.part && !VB) {
        // the slice's partial vector, read once by the merge: non-temporal
        double *ps = a.part + size_t(s) * a.nrows + row0;
        for (uint32_t k = tid; k < rows; k += nt) __builtin_nontemporal_store(yl[k], ps + k);
    } else {
        for (uint32_t k = tid; k < rows; k += nt) unsafeAtomicAdd(out + k, yl[k]);
    }
}

// Value gradient in entry order: vbar[e] = ybar[row(e)] * x[col(e)].  The
// tile's ybar rows sit in LDS; x is gathered in the tile's column order;
// vbar goes out with non-temporal stores.
template <int KEY>
__global__ void __launch_bounds__(1024) rct_vbar_kernel(RctArgs a, const double *__restrict__ ybar,
                                                        const double *__restrict__ x, double *__restrict__ vbar) {
    extern __shared__ double yb[];
    TileOf tile;
    if (!tile.from(a, blockIdx.x)) return;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t row0 = a.rb_start[tile.rb], rows = a.rb_start[tile.rb + 1] - row0;
    for (uint32_t k = tid; k < rows; k += nt) yb[k] = ybar[row0 + k];
    __syncthreads();
    const uint32_t t = tile.rb * a.S + tile.s;
    const uint32_t beg = a.tile_ptr[t], end = a.tile_ptr[t + 1];
    const RctEntries<double, KEY> ent(a, row0);  // columns and rows only
    constexpr int U = 4;
    uint32_t e = beg + tid;
    for (; e + (U - 1) * nt < end; e += U * nt) {
        uint32_t c[U], r[U];
#pragma unroll
        for (int j = 0; j < U; ++j) ent.at(e + j * nt, c[j], r[j]);
        double xv[U];
#pragma unroll
        for (int j = 0; j < U; ++j) xv[j] = x[c[j]];
#pragma unroll
        for (int j = 0; j < U; ++j) __builtin_nontemporal_store(yb[r[j]] * xv[j], vbar + e + j * nt);
    }
    for (; e < end; e += nt) {
        uint32_t c, r;
        ent.at(e, c, r);
        __builtin_nontemporal_store(yb[r] * x[c], vbar + e);
    }
}

// SpMM: Y[:, k0 .. k0 + KT) (+)= A X[:, k0 .. k0 + KT) over one tile.  KT
// lanes serve one entry (one coalesced KT x 8-byte gather of the entry's X
// row); the tile's rows x KT sums stay in LDS (ds_add_f64) and go to Y at
// the end: stored with one slice, added with FP64 atomics (Y cleared first)
// with several.
template <class VT, int KEY, int KT>
__global__ void __launch_bounds__(1024) rct_spmm_kernel(RctArgs a, uint32_t k, uint32_t k0,
                                                        const double *__restrict__ X, double *__restrict__ Y) {
    extern __shared__ double acc[];
    TileOf tile;
    if (!tile.from(a, blockIdx.x)) return;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t row0 = a.rb_start[tile.rb], rows = a.rb_start[tile.rb + 1] - row0;
    for (uint32_t q = tid; q < rows * KT; q += nt) acc[q] = 0.0;
    __syncthreads();
    const uint32_t t = tile.rb * a.S + tile.s;
    const uint32_t beg = a.tile_ptr[t], end = a.tile_ptr[t + 1];
    const RctEntries<VT, KEY> ent(a, row0);
    const uint32_t j = tid % KT, nsub = nt / KT;  // nt is a multiple of 64
    const bool live = k0 + j < k;                 // the last k-tile may be partial
    const double *__restrict__ xj = X + k0 + j;
    constexpr int U = 4;
    uint32_t e = beg + tid / KT;
    for (; e + (U - 1) * nsub < end; e += U * nsub) {
        uint32_t c[U], r[U];
        double v[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ent.load(e + u * nsub, c[u], r[u], v[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = live ? xj[size_t(c[u]) * k] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) atomicAdd(&acc[r[u] * KT + j], v[u] * xv[u]);
    }
    for (; e < end; e += nsub) {
        uint32_t c, r;
        double v;
        ent.load(e, c, r, v);
        const double xv = live ? xj[size_t(c) * k] : 0.0;
        atomicAdd(&acc[r * KT + j], v * xv);
    }
    __syncthreads();
    for (uint32_t q = tid; q < rows * KT; q += nt) {
        const uint32_t rr = q / KT, jj = q % KT;
        if (k0 + jj >= k) continue;
        double *out = Y + size_t(row0 + rr) * k + k0 + jj;
        if (a.S == 1)
            *out = acc[q];
        else
            unsafeAtomicAdd(out, acc[q]);
    }
}

template <class VT, int KEY>
hipError_t rct_spmm_vt(const RctArgs &a, uint32_t k, uint32_t kt, const double *X, double *Y, hipStream_t s) {
    const size_t lds = size_t(a.R) * kt * sizeof(double);
    const uint32_t per = 8 / a.S;
    const unsigned grid = (a.nrb + per - 1) / per * 8;
    for (uint32_t k0 = 0; k0 < k; k0 += kt) {
        switch (kt) {
            case 8: rct_spmm_kernel<VT, KEY, 8><<<grid, a.block, lds, s>>>(a, k, k0, X, Y); break;
            case 16: rct_spmm_kernel<VT, KEY, 16><<<grid, a.block, lds, s>>>(a, k, k0, X, Y); break;
            case 32: rct_spmm_kernel<VT, KEY, 32><<<grid, a.block, lds, s>>>(a, k, k0, X, Y); break;
            case 64: rct_spmm_kernel<VT, KEY, 64><<<grid, a.block, lds, s>>>(a, k, k0, X, Y); break;
            default: return hipErrorInvalidValue;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t rct_spmm(const RctArgs &a, uint32_t k, uint32_t kt, const double *X, double *Y, hipStream_t s) {
    if (!a.nrows || !k) return hipSuccess;
    if (a.S > 1) {
        const hipError_t e = clear_f64(Y, size_t(a.nrows) * k, s);
        if (e != hipSuccess) return e;
    }
    const int km = key_mode(a);
    if (a.vtype == 1)
        return km == KEY_DIAG     ? rct_spmm_vt<float, KEY_DIAG>(a, k, kt, X, Y, s)
               : km == KEY_PACKED ? rct_spmm_vt<float, KEY_PACKED>(a, k, kt, X, Y, s)
                                  : rct_spmm_vt<float, KEY_SPLIT>(a, k, kt, X, Y, s);
    return km == KEY_DIAG     ? rct_spmm_vt<double, KEY_DIAG>(a, k, kt, X, Y, s)
           : km == KEY_PACKED ? rct_spmm_vt<double, KEY_PACKED>(a, k, kt, X, Y, s)
                              : rct_spmm_vt<double, KEY_SPLIT>(a, k, kt, X, Y, s);
}

hipError_t rct_vbar(const RctArgs &a, const double *ybar, const double *x, double *vbar, hipStream_t s) {
    if (!a.nrows) return hipSuccess;
    const uint32_t per = 8 / a.S;
    const unsigned grid = (a.nrb + per - 1) / per * 8;
    const size_t lds = size_t(a.R) * sizeof(double);
    switch (key_mode(a)) {
        case KEY_DIAG: rct_vbar_kernel<KEY_DIAG><<<grid, a.block, lds, s>>>(a, ybar, x, vbar); break;
        case KEY_PACKED: rct_vbar_kernel<KEY_PACKED><<<grid, a.block, lds, s>>>(a, ybar, x, vbar); break;
        default: rct_vbar_kernel<KEY_SPLIT><<<grid, a.block, lds, s>>>(a, ybar, x, vbar);
    }
    return hipGetLastError();
}

hipError_t rct(const RctArgs &a, const double *x, double *y, hipStream_t s, hipEvent_t mid, const double *xr,
               double *vbar, const double *rb, double *res) {
    if (!a.nrows) return hipSuccess;
    if (vbar) rb = res = nullptr;  // the residual belongs to a forward pass
    const uint32_t per = 8 / a.S;
    const uint32_t groups = (a.nrb + per - 1) / per;
    const unsigned grid = groups * 8;
    const bool vb = vbar != nullptr;
    const size_t lds = size_t(a.R) * sizeof(double) * (vb ? 2 : 1);
    const bool partials = a.S > 1 && a.part && !vb;
    if (a.S > 1 && !partials) {
        const hipError_t e = clear_f64(y, a.nrows, s);
        if (e != hipSuccess) return e;
        if (mid) {
            const hipError_t e = hipEventRecord(mid, s);
            if (e != hipSuccess) return e;
        }
    }
    const int km = key_mode(a);
    // the value-gradient sweep too streams with non-temporal loads where the
    // plan asks for them (its v_bar stores are non-temporal anyway): the
    // permuted 128^3 stencil's forward + adjoint, renumbered, 472.8 -> 467.7
    // us, one box
#define AX_RCT_LAUNCH(VT, P)                                                                         \
    do {                                                                                             \
        if (vb && a.nt_stream)                                                                       \
            rct_kernel<VT, P, true, true><<<grid, a.block, lds, s>>>(a, x, y, xr, vbar, rb, res);   \
        else if (vb)                                                                                 \
            rct_kernel<VT, P, true, false><<<grid, a.block, lds, s>>>(a, x, y, xr, vbar, rb, res);  \
        else if (a.nt_stream)                                                                        \
            rct_kernel<VT, P, false, true><<<grid, a.block, lds, s>>>(a, x, y, xr, vbar, rb, res);  \
        else                                                                                         \
            rct_kernel<VT, P, false, false><<<grid, a.block, lds, s>>>(a, x, y, xr, vbar, rb, res); \
    } while (0)
    if (a.vtype == 1) {
        if (km == KEY_DIAG)
            AX_RCT_LAUNCH(float, KEY_DIAG);
        else if (km == KEY_PACKED)
            AX_RCT_LAUNCH(float, KEY_PACKED);
        else
            AX_RCT_LAUNCH(float, KEY_SPLIT);
    } else {
        if (km == KEY_DIAG)
            AX_RCT_LAUNCH(double, KEY_DIAG);
        else if (km == KEY_PACKED)
            AX_RCT_LAUNCH(double, KEY_PACKED);
        else
            AX_RCT_LAUNCH(double, KEY_SPLIT);
    }
#undef AX_RCT_LAUNCH
    hipError_t e = hipGetLastError();
    if (partials) {
        if (e == hipSuccess && mid) e = hipEventRecord(mid, s);
        if (e == hipSuccess) e = part_merge(a.part, a.S, a.nrows, y, s, rb, res);
    } else if (res && a.S > 1 && e == hipSuccess) {
        e = residual(a.nrows, y, rb, res, s);  // after the slices' adds
    }
    return e;
}

}  // namespace launch
}  // namespace axspmv
