// AXT UNC kernels on row-aligned wave blocks.
//
// The reference runs one FP64 atomicAdd per tile (gaxtuh1hw16w/08w/04w,
// cudaSpmv.cu:1570-1850) or per tile column (gaxtuh, cudaSpmv.cu:1939-1959)
// into a y it zeroed beforehand.  Here every wave owns a contiguous range of
// "units" that starts and ends on row boundaries (built on the host), so each
// row is reduced inside one wave and stored once with a plain store: no
// atomics, no y zeroing, deterministic results.
//
//   unit (th == 1): K consecutive slots of one tile (K divides thw); the lane
//                   reads K values and K column indices with vector loads and
//                   gathers K x entries.
//   unit (th >  1): one tile column g (th slots, one per level); level f of
//                   64 neighbouring columns is contiguous, so the lanes' loads
//                   coalesce.
//
// A wave walks its units 64 at a time.  Lane sums are combined across lanes
// by a DPP segmented scan keyed on the row id (sec[]); the last lane of each
// row segment stores y.  A segment still open at lane 63 is carried to the
// next 64 units in a scalar register.
#include "rowblock.hpp"

namespace {
using namespace axspmv::dev;
using axspmv::launch::Block;

// STAGE: blocks whose window fits STAGE_COLS read x from a per-wave LDS copy
// (banded / stencil structure); the others, and every block when STAGE is
// false, gather from HBM.
template <class Src, bool STAGE>
__global__ void __launch_bounds__(256) rowblock_kernel(Src src, const Block *__restrict__ blocks, uint32_t nblocks,
                                                       const uint32_t *__restrict__ empty, uint32_t nempty,
                                                       const Window *__restrict__ win, const double *__restrict__ x,
                                                       double *__restrict__ y) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wave * 64 + lane; i < nempty; i += nwaves * 64) y[empty[i]] = 0.0;
    if (wave >= nblocks) return;
    const Block b = blocks[wave];
    if constexpr (STAGE) {
        __shared__ double xs[4][axspmv::launch::STAGE_COLS];
        const Window w = win[wave];
        if (w.width <= axspmv::launch::STAGE_COLS) {
            rowblock_process(src, b, stage_window(xs[threadIdx.x >> 6], w, x), y);
            return;
        }
    }
    rowblock_process(src, b, XGlobal{x}, y);
}

// Column-sliced AXT UNC th == 1: each column slice of the matrix is its own
// AXT layout (planes concatenated); the host lists the wave blocks so that
// the four waves of workgroup g all take blocks of slice (g % 8) % S, and
// workgroups go to the 8 XCDs round-robin, so each XCD's L2 holds only its
// slice of x.  A block adds its rows' slice sums into y (cleared first)
// with FP64 atomics: S adds per row at most.
template <class Src>
__global__ void __launch_bounds__(256) rowblock_slices_kernel(Src src, const Block *__restrict__ blocks,
                                                              uint32_t nblocks, const double *__restrict__ x,
                                                              double *__restrict__ y) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= nblocks) return;
    rowblock_process<Src, XGlobal, true>(src, blocks[wave], XGlobal{x}, y);
}

// The deterministic form: slice s's row sums are stored (not added) into
// the partial vector part + s * nrows (s = (g % 8) % S for workgroup g, as
// the dispatch list assigns), rows without entries in a slice get 0 there
// (`empty` lists s * nrows + r), and part_merge adds the S partials of every
// row in slice order.
template <class Src>
__global__ void __launch_bounds__(256) rowblock_partial_kernel(Src src, const Block *__restrict__ blocks,
                                                               uint32_t nblocks, const uint32_t *__restrict__ empty,
                                                               uint32_t nempty, uint32_t S, uint32_t nrows,
                                                               const double *__restrict__ x,
                                                               double *__restrict__ part) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wave * 64 + lane; i < nempty; i += nwaves * 64) part[empty[i]] = 0.0;
    if (wave >= nblocks) return;
    double *ys = part + size_t((blockIdx.x % 8) % S) * nrows;
    rowblock_process<Src, XGlobal, false>(src, blocks[wave], XGlobal{x}, ys);
}

// ------------------------------------------------------------------ adjoint
// Value gradient of the AXT UNC th == 1 layout: plane slot q of row r is CSR
// position rowptr[r] + q - first[r] (first[r] = first plane slot of row r);
// slots past the row length are padding.
using axspmv::launch::VbarArgs;

template <int K>
__device__ __forceinline__ void put_vbar(uint32_t u, uint32_t r, double s, const double *xv, const VbarArgs &va,
                                         double *__restrict__ vbar) {
    const uint32_t off = u * K - va.first[r];
    const uint32_t b = va.rowptr[r], len = va.rowptr[r + 1] - b;
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (off + j < len) vbar[b + off + j] = s * xv[j];
}

// v_bar[k] = y_bar[row(k)] * x[col(k)], one lane per unit of K slots.
template <class VT, int K, bool NT>
__global__ void __launch_bounds__(256) vbar_kernel(UncH1Src<VT, K, NT> src, uint32_t nunits, const double *__restrict__ x,
                                                   const double *__restrict__ ybar, double *__restrict__ vbar,
                                                   VbarArgs va) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const uint32_t r = src.row(u);
    double xv[K];
    src.gather_x(u, x, xv);
    put_vbar<K>(u, r, ybar[r], xv, va, vbar);
}

// Fused forward pass of the least-squares adjoint: y = A x, r = y - b and
// v_bar[k] = r[row(k)] * x[col(k)] in one sweep over A.  Blocks hold at most
// 64 units (one pass of the wave), so each row's total is known while the
// gathered x values are still in registers; the row's residual goes back to
// its lanes with ballot + bpermute.  A block of more than 64 units is a
// single long row and takes a second sweep for v_bar.
template <class VT, int K, bool NT>
__global__ void __launch_bounds__(256) fused_kernel(UncH1Src<VT, K, NT> src, const Block *__restrict__ blocks,
                                                    uint32_t nblocks, const uint32_t *__restrict__ empty,
                                                    uint32_t nempty, const double *__restrict__ x,
                                                    const double *__restrict__ b, double *__restrict__ y,
                                                    double *__restrict__ res, double *__restrict__ vbar, VbarArgs va) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wave * 64 + lane; i < nempty; i += nwaves * 64) {
        const uint32_t r = empty[i];
        y[r] = 0.0;
        res[r] = -b[r];
    }
    if (wave >= nblocks) return;
    const Block bl = blocks[wave];
    if (bl.end - bl.beg <= 64) {
        const uint32_t u = bl.beg + lane;
        const bool act = u < bl.end;
        double xv[K];
        double s = 0.0;
        uint32_t r = NOROW;
        if (act) {
            s = src.sum_keep(u, x, xv);
            r = src.row(u);
        }
        uint32_t head = r != shift_up_u32(r, NOROW);
        const double v = seg_scan_f64(s, head);
        const bool end = act && shift_down_u32(r, NOROW) != r;
        double rv = 0.0;
        if (end) {
            rv = v - b[r];
            y[r] = v;
            res[r] = rv;
        }
        const uint64_t endmask = __ballot(end);
        const int src_lane = act ? int(lane) + __builtin_ctzll(endmask >> lane) : int(lane);
        rv = __shfl(rv, src_lane);
        if (act) put_vbar<K>(u, r, rv, xv, va, vbar);
        return;
    }
    // one long row
    const uint32_t r = src.row(bl.beg);
    double acc = 0.0;
    for (uint32_t u = bl.beg + lane; u < bl.end; u += 64) acc += src.sum(u, XGlobal{x});
    const double tot = group_sum_f64<64>(acc);
    const double rv = tot - b[r];
    if (lane == 0) {
        y[r] = tot;
        res[r] = rv;
    }
    for (uint32_t u = bl.beg + lane; u < bl.end; u += 64) {
        double xv[K];
        src.gather_x(u, x, xv);
        put_vbar<K>(u, r, rv, xv, va, vbar);
    }
}

template <class Src>
hipError_t run(const Src &src, const axspmv::launch::SegArgs &a, const double *x, double *y, hipStream_t s) {
    const uint32_t block = 256;  // 4 waves: the LDS windows are indexed by wave
    uint32_t waves = a.nblocks ? a.nblocks : 1;
    const unsigned grid = (waves + block / 64 - 1) / (block / 64);
    if (a.part) {  // column slices, deterministic: partials, then their sum in slice order
        rowblock_partial_kernel<Src><<<grid, block, 0, s>>>(src, a.blocks, a.nblocks, a.empty, a.nempty, a.S,
                                                            a.nrows_add, x, a.part);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && a.mid) e = hipEventRecord(a.mid, s);
        if (e == hipSuccess) e = axspmv::launch::part_merge(a.part, a.S, a.nrows_add, y, s);
        return e;
    }
    if (a.nrows_add) {  // column slices: clear y, then add every slice's row sums
        const hipError_t c = axspmv::launch::clear_f64(y, a.nrows_add, s);
        if (c != hipSuccess) return c;
        if (a.mid) {
            const hipError_t e = hipEventRecord(a.mid, s);
            if (e != hipSuccess) return e;
        }
        rowblock_slices_kernel<Src><<<grid, block, 0, s>>>(src, a.blocks, a.nblocks, x, y);
        return hipGetLastError();
    }
    if (a.win)
        rowblock_kernel<Src, true><<<grid, block, 0, s>>>(src, a.blocks, a.nblocks, a.empty, a.nempty, a.win, x, y);
    else
        rowblock_kernel<Src, false><<<grid, block, 0, s>>>(src, a.blocks, a.nblocks, a.empty, a.nempty, nullptr, x, y);
    return hipGetLastError();
}

template <class VT, bool NT>
hipError_t unc_h1_vt(const axspmv::launch::SegArgs &a, const double *x, double *y, hipStream_t s) {
    const uint32_t thw_log = __builtin_ctz(a.thw);
    const VT *vp = static_cast<const VT *>(a.vplane);
    if (a.thw >= 4) return run(UncH1Src<VT, 4, NT>{vp, a.cplane, a.rowsrc, thw_log - 2}, a, x, y, s);
    if (a.thw == 2) return run(UncH1Src<VT, 2, NT>{vp, a.cplane, a.rowsrc, 0}, a, x, y, s);
    return run(UncH1Src<VT, 1, NT>{vp, a.cplane, a.rowsrc, 0}, a, x, y, s);
}

template <class VT, bool NT>
hipError_t unc_h_vt(const axspmv::launch::SegArgs &a, const double *x, double *y, hipStream_t s) {
    const uint32_t thw_log = __builtin_ctz(a.thw);
    const VT *vp = static_cast<const VT *>(a.vplane);
    switch (a.th) {
#define AX_TH_CASE(T) \
    case T: return run(UncHSrc<VT, T, NT>{vp, a.cplane, a.rowsrc, thw_log, a.th}, a, x, y, s);
        AX_TH_CASE(2)
        AX_TH_CASE(4)
        AX_TH_CASE(6)
        AX_TH_CASE(8)
        AX_TH_CASE(10)
        AX_TH_CASE(12)
        AX_TH_CASE(16)
#undef AX_TH_CASE
        default: return run(UncHSrc<VT, 0, NT>{vp, a.cplane, a.rowsrc, thw_log, a.th}, a, x, y, s);
    }
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t axt_unc_h1(const SegArgs &a, const double *x, double *y, hipStream_t s) {
    if (a.nt) return a.vtype == 1 ? unc_h1_vt<float, true>(a, x, y, s) : unc_h1_vt<double, true>(a, x, y, s);
    return a.vtype == 1 ? unc_h1_vt<float, false>(a, x, y, s) : unc_h1_vt<double, false>(a, x, y, s);
}

hipError_t axt_unc_h(const SegArgs &a, const double *x, double *y, hipStream_t s) {
    if (a.nt) return a.vtype == 1 ? unc_h_vt<float, true>(a, x, y, s) : unc_h_vt<double, true>(a, x, y, s);
    return a.vtype == 1 ? unc_h_vt<float, false>(a, x, y, s) : unc_h_vt<double, false>(a, x, y, s);
}

hipError_t axt_unc_h1_ax(const SegArgs &a, double *y, hipStream_t s) {
    const uint32_t thw_log = __builtin_ctz(a.thw);
    const double *ax = static_cast<const double *>(a.vplane);
    if (a.thw >= 4) return run(AxH1Src<4>{ax, a.rowsrc, thw_log - 2, thw_log}, a, nullptr, y, s);
    if (a.thw == 2) return run(AxH1Src<2>{ax, a.rowsrc, 0, thw_log}, a, nullptr, y, s);
    return run(AxH1Src<1>{ax, a.rowsrc, 0, thw_log}, a, nullptr, y, s);
}

namespace {
template <class VT, int K, bool NT>
hipError_t vbar_vk(const SegArgs &a, uint32_t ushift, uint32_t nunits, const double *x, const double *ybar,
                   double *vbar, const VbarArgs &va, hipStream_t s) {
    const UncH1Src<VT, K, NT> src{static_cast<const VT *>(a.vplane), a.cplane, a.rowsrc, ushift};
    if (nunits) vbar_kernel<VT, K, NT><<<(nunits + 255) / 256, 256, 0, s>>>(src, nunits, x, ybar, vbar, va);
    return hipGetLastError();
}
template <class VT, int K, bool NT>
hipError_t fused_vk(const SegArgs &a, uint32_t ushift, const double *x, const double *b, double *y, double *res,
                    double *vbar, const VbarArgs &va, hipStream_t s) {
    const UncH1Src<VT, K, NT> src{static_cast<const VT *>(a.vplane), a.cplane, a.rowsrc, ushift};
    const uint32_t waves = a.nblocks ? a.nblocks : 1;
    fused_kernel<VT, K, NT><<<(waves + 3) / 4, 256, 0, s>>>(src, a.blocks, a.nblocks, a.empty, a.nempty, x, b, y,
                                                           res, vbar, va);
    return hipGetLastError();
}
template <class VT>
hipError_t vbar_v(const SegArgs &a, uint32_t nunits, const double *x, const double *ybar, double *vbar,
                  const VbarArgs &va, hipStream_t s) {
    const uint32_t tl = __builtin_ctz(a.thw);
    if (a.nt) {
        if (a.thw >= 4) return vbar_vk<VT, 4, true>(a, tl - 2, nunits, x, ybar, vbar, va, s);
        if (a.thw == 2) return vbar_vk<VT, 2, true>(a, 0, nunits, x, ybar, vbar, va, s);
        return vbar_vk<VT, 1, true>(a, 0, nunits, x, ybar, vbar, va, s);
    }
    if (a.thw >= 4) return vbar_vk<VT, 4, false>(a, tl - 2, nunits, x, ybar, vbar, va, s);
    if (a.thw == 2) return vbar_vk<VT, 2, false>(a, 0, nunits, x, ybar, vbar, va, s);
    return vbar_vk<VT, 1, false>(a, 0, nunits, x, ybar, vbar, va, s);
}
// The plan's non-temporal plane loads (a.nt) apply here too: forward +
// adjoint on the 2^20-row power-law matrix 548.8 -> 531.0 us fused, 738.9
// -> 733.0 us split (one box).
template <class VT>
hipError_t fused_v(const SegArgs &a, const double *x, const double *b, double *y, double *res, double *vbar,
                   const VbarArgs &va, hipStream_t s) {
    const uint32_t tl = __builtin_ctz(a.thw);
    if (a.nt) {
        if (a.thw >= 4) return fused_vk<VT, 4, true>(a, tl - 2, x, b, y, res, vbar, va, s);
        if (a.thw == 2) return fused_vk<VT, 2, true>(a, 0, x, b, y, res, vbar, va, s);
        return fused_vk<VT, 1, true>(a, 0, x, b, y, res, vbar, va, s);
    }
    if (a.thw >= 4) return fused_vk<VT, 4, false>(a, tl - 2, x, b, y, res, vbar, va, s);
    if (a.thw == 2) return fused_vk<VT, 2, false>(a, 0, x, b, y, res, vbar, va, s);
    return fused_vk<VT, 1, false>(a, 0, x, b, y, res, vbar, va, s);
}
}  // namespace

hipError_t axt_unc_h1_vbar(const SegArgs &a, uint32_t nunits, const double *x, const double *ybar, double *vbar,
                           const VbarArgs &va, hipStream_t s) {
    return a.vtype == 1 ? vbar_v<float>(a, nunits, x, ybar, vbar, va, s) : vbar_v<double>(a, nunits, x, ybar, vbar, va, s);
}

hipError_t axt_unc_h1_fused(const SegArgs &a, const double *x, const double *b, double *y, double *res,
                            double *vbar, const VbarArgs &va, hipStream_t s) {
    return a.vtype == 1 ? fused_v<float>(a, x, b, y, res, vbar, va, s) : fused_v<double>(a, x, b, y, res, vbar, va, s);
}

}  // namespace launch
}  // namespace axspmv
