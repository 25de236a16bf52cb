// RCT: row-block x column-slice tiles with entries sorted by column.
//
// Random column indices make SpMV a gather problem: every nonzero costs one
// L2 request for x, and once x outgrows an XCD's 4 MB L2 the requests go to
// the Infinity Cache.  RCT attacks both:
//
//  * the matrix is cut into tiles of R rows x C columns (S = ncols / C column
//    slices); a workgroup owns one tile and accumulates its R rows in LDS;
//  * entries of a tile are stored sorted by column, so neighbouring lanes
//    gather neighbouring x entries -- one cache line serves several lanes;
//  * workgroups are dispatched round-robin to the 8 XCDs (workgroup w runs on
//    XCD w % 8), and tile w belongs to column slice (w % 8) % S: each XCD only
//    ever reads its own slice of x, which stays resident in its L2.
//
// Each tile writes its R partial sums (zero for rows it does not touch) to
// the partial vector of its slice; a second kernel adds the S partials of
// every row in slice order.  The LDS adds (ds_add_f64) make the order of the
// additions inside a row nondeterministic; the result stays within the FP64
// tolerance of the left-to-right CPU sum.
#include "launch.hpp"
#include "wave64.hpp"

namespace {
using axspmv::launch::RctArgs;

template <class VT>
__global__ void __launch_bounds__(1024) rct_kernel(RctArgs a, const double *__restrict__ x) {
    extern __shared__ double yl[];
    const uint32_t w = blockIdx.x;
    const uint32_t xcd = w % 8, i = w / 8;
    const uint32_t per = 8 / a.S;  // row-block groups per slice
    const uint32_t s = xcd % a.S, rb = i * per + xcd / a.S;
    if (rb >= a.nrb) return;  // the whole workgroup leaves together
    const uint32_t R = a.R, tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t k = tid; k < R; k += nt) yl[k] = 0.0;
    __syncthreads();
    const uint32_t t = rb * a.S + s;  // tile index in storage order
    const uint32_t beg = a.tile_ptr[t], end = a.tile_ptr[t + 1];
    const VT *__restrict__ val = static_cast<const VT *>(a.val);
    uint32_t e = beg + tid;
    for (; e + 3 * nt < end; e += 4 * nt) {
        const uint32_t e1 = e + nt, e2 = e + 2 * nt, e3 = e + 3 * nt;
        const uint32_t c0 = a.col[e], c1 = a.col[e1], c2 = a.col[e2], c3 = a.col[e3];
        const uint32_t r0 = a.lrow[e], r1 = a.lrow[e1], r2 = a.lrow[e2], r3 = a.lrow[e3];
        const double v0 = double(val[e]) * x[c0], v1 = double(val[e1]) * x[c1];
        const double v2 = double(val[e2]) * x[c2], v3 = double(val[e3]) * x[c3];
        atomicAdd(&yl[r0], v0);
        atomicAdd(&yl[r1], v1);
        atomicAdd(&yl[r2], v2);
        atomicAdd(&yl[r3], v3);
    }
    for (; e < end; e += nt) atomicAdd(&yl[a.lrow[e]], double(val[e]) * x[a.col[e]]);
    __syncthreads();
    const uint32_t row0 = rb * R;
    const uint32_t rows = min(R, a.nrows - row0);
    double *out = a.part + size_t(s) * a.nrows + row0;
    for (uint32_t k = tid; k < rows; k += nt) out[k] = yl[k];
}

// y[r] = sum over slices s of part[s][r], in slice order.
__global__ void __launch_bounds__(256) rct_merge(const double *__restrict__ part, uint32_t S, uint32_t n,
                                                 double *__restrict__ y) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double acc = part[r];
    for (uint32_t s = 1; s < S; ++s) acc += part[size_t(s) * n + r];
    y[r] = acc;
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t rct(const RctArgs &a, const double *x, double *y, hipStream_t s) {
    if (!a.nrows) return hipSuccess;
    const uint32_t per = 8 / a.S;
    const uint32_t groups = (a.nrb + per - 1) / per;
    const unsigned grid = groups * 8;
    const size_t lds = size_t(a.R) * sizeof(double);
    // With S == 1 the tile's sums are the rows' totals: write y directly.
    RctArgs b = a;
    if (a.S == 1) b.part = y;
    if (a.vtype == 1)
        rct_kernel<float><<<grid, a.block, lds, s>>>(b, x);
    else
        rct_kernel<double><<<grid, a.block, lds, s>>>(b, x);
    if (a.S > 1) rct_merge<<<(a.nrows + 255) / 256, 256, 0, s>>>(a.part, a.S, a.nrows, y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace axspmv
