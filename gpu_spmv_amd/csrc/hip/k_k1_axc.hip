// K1 and AXC kernels.
//
// K1 (reference gk1, cudaSpmv.cu:851-867): a chunk of C rows stored column
// major (element j of row r of chunk c at chp[c] + j*C + r), one lane per
// row, so every step of the 64 lanes reads 64 neighbouring values.  Here a
// 64-row chunk is one wavefront.  Rows sorted by length put the longest rows
// together; a chunk whose nmc exceeds the per-wave step limit is split over
// several waves (work items) whose partial sums a second kernel adds in a
// fixed order -- the reference instead runs one thread for the whole row.
// The reference's first step is unconditional and reads past an all-empty
// chunk (nmc = 0); here the loop simply does not run.
//
// AXC (reference gaxc, cudaSpmv.cu:1072-1096): each row is a run of bricks
// of hbs slots (value plane slot brp[r]/2 + k, column in mapx).  A group of
// G = hbs lanes walks one row brick by brick and reduces with DPP; rows
// longer than 64 bricks are taken by a whole wave (listed on the host), so a
// few long rows do not serialise on G lanes.
#include "launch.hpp"
#include "wave64.hpp"

namespace {
using namespace axspmv::dev;
using axspmv::launch::K1Item;

// NT: val and col are streamed with non-temporal loads.
template <class VT, bool NT>
__global__ void __launch_bounds__(256) k1_kernel(const K1Item *__restrict__ items, uint32_t nitems, uint32_t chunk,
                                                 const VT *__restrict__ val, const uint32_t *__restrict__ col,
                                                 const uint32_t *__restrict__ permi, uint32_t nrows,
                                                 const double *__restrict__ x, double *__restrict__ y,
                                                 double *__restrict__ part) {
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t per_wave = 64 / chunk;  // chunks per wave (1 for chunk 64)
    const uint32_t wave = gl >> 6, lane = gl & 63;
    const uint32_t it = wave * per_wave + lane / chunk;
    if (it >= nitems) return;
    const K1Item w = items[it];
    const uint32_t r = lane % chunk;
    const uint32_t gid = w.chunk * chunk + r;
    const uint64_t base = uint64_t(w.chp) + r;
    double acc = 0.0;
    uint32_t j = w.j0;
    for (; j + 4 <= w.j1; j += 4) {
        const uint64_t q = base + uint64_t(j) * chunk;
        const VT v0 = stream_load<NT>(val + q), v1 = stream_load<NT>(val + q + chunk),
                 v2 = stream_load<NT>(val + q + 2 * chunk), v3 = stream_load<NT>(val + q + 3 * chunk);
        const double x0 = xpad(x, stream_load<NT>(col + q)), x1 = xpad(x, stream_load<NT>(col + q + chunk)),
                     x2 = xpad(x, stream_load<NT>(col + q + 2 * chunk)),
                     x3 = xpad(x, stream_load<NT>(col + q + 3 * chunk));
        acc = fma(double(v0), x0, acc);
        acc = fma(double(v1), x1, acc);
        acc = fma(double(v2), x2, acc);
        acc = fma(double(v3), x3, acc);
    }
    for (; j < w.j1; ++j) {
        const uint64_t q = base + uint64_t(j) * chunk;
        acc = fma(double(stream_load<NT>(val + q)), xpad(x, stream_load<NT>(col + q)), acc);
    }
    if (w.slot == NOROW) {
        if (gid < nrows) y[permi[gid]] = acc;
    } else {
        part[uint64_t(w.slot) * chunk + r] = acc;
    }
}

// Sum of the partial results of one split chunk, in item order.
__global__ void __launch_bounds__(256) k1_fixup(const uint32_t *__restrict__ split, uint32_t nsplit, uint32_t chunk,
                                                const uint32_t *__restrict__ permi, uint32_t nrows,
                                                const double *__restrict__ part, double *__restrict__ y) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = g / chunk, r = g % chunk;
    if (s >= nsplit) return;
    const uint32_t c = split[3 * s], slot0 = split[3 * s + 1], ns = split[3 * s + 2];
    const uint32_t gid = c * chunk + r;
    double acc = 0.0;
    for (uint32_t k = 0; k < ns; ++k) acc += part[uint64_t(slot0 + k) * chunk + r];
    if (gid < nrows) y[permi[gid]] = acc;
}

// AXC: G lanes per row; rows listed in `longrows` are skipped here and done
// by axc_long (one wave per row).
template <class VT, int G, bool NT>
__global__ void __launch_bounds__(256) axc_kernel(uint32_t nrows, const uint32_t *__restrict__ brp,
                                                  const VT *__restrict__ vp, const uint32_t *__restrict__ mapx,
                                                  uint32_t long_slots, const double *__restrict__ x,
                                                  double *__restrict__ y) {
    const size_t gl = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gl / G >= nrows) return;  // whole groups leave together (G divides 64)
    const uint32_t r = uint32_t(gl / G), g = uint32_t(gl % G);
    const uint32_t b = brp[r] >> 1, e = brp[r + 1] >> 1;
    if (e - b > long_slots) return;
    double acc = 0.0;
    for (uint32_t q = b + g; q < e; q += G) acc = fma(double(stream_load<NT>(vp + q)), xpad(x, stream_load<NT>(mapx + q)), acc);
    acc = group_sum_f64<G>(acc);
    if (g == 0) y[r] = acc;
}

template <class VT, bool NT>
__global__ void __launch_bounds__(256) axc_long(const uint32_t *__restrict__ rows, uint32_t nlong,
                                                const uint32_t *__restrict__ brp, const VT *__restrict__ vp,
                                                const uint32_t *__restrict__ mapx, const double *__restrict__ x,
                                                double *__restrict__ y) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= nlong) return;
    const uint32_t r = rows[wave];
    const uint32_t b = brp[r] >> 1, e = brp[r + 1] >> 1;
    double acc = 0.0;
    for (uint32_t q = b + lane; q < e; q += 64)
        acc = fma(double(stream_load<NT>(vp + q)), xpad(x, stream_load<NT>(mapx + q)), acc);
    acc = group_sum_f64<64>(acc);
    if (lane == 0) y[r] = acc;
}

template <class VT>
hipError_t k1_vt(const axspmv::launch::K1Args &a, const double *x, double *y, hipStream_t s) {
    const uint32_t per_wave = 64 / a.chunk;
    const uint32_t waves = (a.nitems + per_wave - 1) / per_wave;
    if (waves && a.nt)
        k1_kernel<VT, true><<<(waves + 3) / 4, 256, 0, s>>>(a.items, a.nitems, a.chunk, static_cast<const VT *>(a.val),
                                                            a.col, a.permi, a.nrows, x, y, a.part);
    else if (waves)
        k1_kernel<VT, false><<<(waves + 3) / 4, 256, 0, s>>>(a.items, a.nitems, a.chunk, static_cast<const VT *>(a.val),
                                                             a.col, a.permi, a.nrows, x, y, a.part);
    if (a.nsplit) {
        const uint32_t th = a.nsplit * a.chunk;
        k1_fixup<<<(th + 255) / 256, 256, 0, s>>>(a.split, a.nsplit, a.chunk, a.permi, a.nrows, a.part, y);
    }
    return hipGetLastError();
}

template <class VT, bool NT>
hipError_t axc_vt(const axspmv::launch::AxcArgs &a, const double *x, double *y, hipStream_t s) {
    const VT *vp = static_cast<const VT *>(a.vplane);
    const uint64_t th = uint64_t(a.nrows) * a.hbs;
    const unsigned grid = unsigned((th + 255) / 256);
    if (a.nrows) switch (a.hbs) {
#define AX_G(G) \
    case G: axc_kernel<VT, G, NT><<<grid, 256, 0, s>>>(a.nrows, a.brp, vp, a.mapx, a.long_slots, x, y); break;
        AX_G(1)
        AX_G(2)
        AX_G(4)
        AX_G(8)
        AX_G(16)
        AX_G(32)
        AX_G(64)
#undef AX_G
        default: return hipErrorInvalidValue;
    }
    if (a.nlong) axc_long<VT, NT><<<(a.nlong + 3) / 4, 256, 0, s>>>(a.longrows, a.nlong, a.brp, vp, a.mapx, x, y);
    return hipGetLastError();
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t k1(const K1Args &a, const double *x, double *y, hipStream_t s) {
    return a.vtype == 1 ? k1_vt<float>(a, x, y, s) : k1_vt<double>(a, x, y, s);
}

hipError_t axc(const AxcArgs &a, const double *x, double *y, hipStream_t s) {
    if (a.nt) return a.vtype == 1 ? axc_vt<float, true>(a, x, y, s) : axc_vt<double, true>(a, x, y, s);
    return a.vtype == 1 ? axc_vt<float, false>(a, x, y, s) : axc_vt<double, false>(a, x, y, s);
}

}  // namespace launch
}  // namespace axspmv
