// Batched SpMV: many matrices of different sizes in one persistent launch.
//
// Every matrix is stored in the AXT UNC layout its own tile parameters call
// for (tile half width / height chosen per matrix), all planes concatenated
// in one device buffer; x and y of the batch are concatenated too.  The host
// cuts every matrix into row-aligned wave blocks and lists all blocks of the
// batch longest first.  A persistent grid (a fixed number of waves sized to
// the chip) walks that list with a grid stride -- every wave leaves after
// its last item -- and runs the same row-block engine as the single-matrix
// kernels on each item.
#include "rowblock.hpp"

namespace {
using namespace axspmv::dev;
using axspmv::launch::BatchArgs;
using axspmv::launch::BatchItem;
using axspmv::launch::BatchMat;

template <class Src>
__device__ __forceinline__ void batch_item(const Src &src, const Block b, const double *__restrict__ xm,
                                           double *__restrict__ ym, const Window *win, uint32_t it, double *xw) {
    if (xw) {
        const Window w = win[it];
        if (w.width <= axspmv::launch::STAGE_COLS) {
            rowblock_process(src, b, stage_window(xw, w, xm), ym);
            return;
        }
    }
    rowblock_process(src, b, XGlobal{xm}, ym);
}

// STAGE: items whose column window fits STAGE_COLS read x from the wave's LDS
// copy (window columns are relative to the item's matrix).
template <class VT, bool STAGE>
__global__ void __launch_bounds__(256) batch_kernel(BatchArgs a, const double *__restrict__ x, double *__restrict__ y) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wave * 64 + lane; i < a.nempty; i += nwaves * 64) y[a.empty[i]] = 0.0;
    const VT *vbase = static_cast<const VT *>(a.vplane);
    double *xw = nullptr;
    if constexpr (STAGE) {
        __shared__ double xs[4][axspmv::launch::STAGE_COLS];
        xw = xs[threadIdx.x >> 6];
    }
    for (uint32_t it = wave; it < a.nitems; it += nwaves) {
        const BatchItem w = a.items[it];
        const BatchMat m = a.mats[w.mat];
        const VT *vp = vbase + m.voff;
        const uint32_t *cp = a.cplane + m.voff;
        const uint32_t *sec = a.sec + m.soff;
        const double *xm = x + m.xoff;
        double *ym = y + m.yoff;
        const Block b{w.beg, w.end};
        switch (m.kind) {
            case 0: batch_item(UncH1Src<VT, 4>{vp, cp, sec, m.ushift}, b, xm, ym, a.win, it, xw); break;
            case 1: batch_item(UncH1Src<VT, 2>{vp, cp, sec, 0}, b, xm, ym, a.win, it, xw); break;
            case 2: batch_item(UncH1Src<VT, 1>{vp, cp, sec, 0}, b, xm, ym, a.win, it, xw); break;
            default: batch_item(UncHSrc<VT, 0>{vp, cp, sec, m.thw_log, m.th}, b, xm, ym, a.win, it, xw); break;
        }
    }
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t batch(const BatchArgs &a, const double *x, double *y, hipStream_t s) {
    const uint32_t waves = a.nwaves ? a.nwaves : 1;
    const unsigned grid = (waves + 3) / 4;
    if (a.vtype == 1) {
        if (a.win)
            batch_kernel<float, true><<<grid, 256, 0, s>>>(a, x, y);
        else
            batch_kernel<float, false><<<grid, 256, 0, s>>>(a, x, y);
    } else {
        if (a.win)
            batch_kernel<double, true><<<grid, 256, 0, s>>>(a, x, y);
        else
            batch_kernel<double, false><<<grid, 256, 0, s>>>(a, x, y);
    }
    return hipGetLastError();
}

}  // namespace launch
}  // namespace axspmv
