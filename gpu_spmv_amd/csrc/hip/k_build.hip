// Format builders on the device: CSR (device arrays) -> AXC / AXT arrays in
// the reference's layouts, equal bit for bit to the host builders.
//
// Each layout is a count, an exclusive scan and a fill:
//   count : per row, the units it occupies (AXC / AXT UNC th == 1: tiles of
//           hw slots, ceil(rl / hw); AXT UNC th > 1: tile columns,
//           ceil(rl / th); COM layouts are packed by CSR position and need
//           no count);
//   scan  : rocPRIM's device exclusive scan over the counts (64-bit, so the
//           host can reject layouts whose lengths overflow 32 bits);
//   fill  : a group of 8 lanes per row writes the row's values and x values
//           into ax, its columns into the column plane (AXC: mapx) and the
//           row / segment headers into sec, and the row's start into con.
// Padding stays as the zero fill of the buffers, as the reference's calloc.
#include <rocprim/device/device_scan.hpp>

#include "launch.hpp"

namespace {
using axspmv::launch::BuildGeo;
using axspmv::launch::BuildOut;
using axspmv::launch::DevCsr;

constexpr uint32_t GROUP = 8;  // lanes per row in the fill

// The device CSR's offsets and columns, checked before the fill reads col /
// val through them: bit 1 offsets decrease, 2 row[0] != 0, 4 row[nrows] !=
// nnz, 8 a column >= ncols.  One thread per row and per entry; reads only
// row[0 .. nrows] and col[0 .. nnz).
__global__ void __launch_bounds__(256) check_kernel(DevCsr a, uint32_t *__restrict__ flag) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t bad = 0;
    if (i < a.nrows && a.row[i + 1] < a.row[i]) bad |= 1u;
    if (i == 0 && a.row[0] != 0) bad |= 2u;
    if (i == 0 && a.row[a.nrows] != a.nnz) bad |= 4u;
    if (i < a.nnz && a.col[i] >= a.ncols) bad |= 8u;
    if (bad) atomicOr(flag, bad);
}

__global__ void __launch_bounds__(256) count_kernel(DevCsr a, uint32_t div, uint64_t *__restrict__ cnt) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.nrows) return;
    const uint32_t rl = a.row[r + 1] - a.row[r];
    cnt[r] = (rl + uint64_t(div) - 1) / div;
}

// Plane slot q -> ax index: groups of hw values followed by hw x values.
__device__ __forceinline__ uint64_t ax_of_plane(uint64_t q, uint32_t hw) { return (q / hw) * 2 * hw + q % hw; }

// A column outside the matrix (an invalid CSR) reads no x and leaves 0.
__device__ __forceinline__ void put(const BuildOut &o, uint32_t hw, uint64_t q, uint32_t col, double val,
                                    const double *__restrict__ x, uint32_t ncols) {
    const uint64_t s = ax_of_plane(q, hw);
    o.ax[s] = val;
    if (x && col < ncols) o.ax[s + hw] = x[col];
    o.plane[q] = col;
}

// KIND 0: AXC / AXT UNC th == 1; 1: AXT UNC th > 1; 2: COM th == 1; 3: COM th > 1.
template <int KIND>
__global__ void __launch_bounds__(256) fill_kernel(DevCsr a, const uint64_t *__restrict__ start, BuildGeo g,
                                                   const double *__restrict__ x, BuildOut o) {
    const uint64_t gid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t r = uint32_t(gid / GROUP), j = uint32_t(gid % GROUP);
    if (r >= a.nrows) return;
    const uint32_t b = a.row[r], l = a.row[r + 1] - b;
    const uint32_t hw = g.hw, th = g.th;
    const uint64_t ths = uint64_t(hw) * th;
    if constexpr (KIND == 0) {
        const uint64_t s = start[r] * hw;  // first plane slot of the row
        if (j == 0) {
            o.con[r + 1] = uint32_t(start[r + 1] * 2 * hw);
            if (r == 0) o.con[0] = 0;
        }
        for (uint32_t k = j; k < l; k += GROUP) {
            put(o, hw, s + k, a.col[b + k], a.val[b + k], x, a.ncols);
            if (o.sec && k % hw == 0) o.sec[(s + k) / hw] = r;  // AXC has no sec
        }
    } else if constexpr (KIND == 1) {
        const uint64_t c0 = start[r];  // first global tile column of the row
        if (j == 0) o.con[r] = uint32_t((c0 / hw) * 2 * ths + c0 % hw);
        for (uint32_t k = j; k < l; k += GROUP) {
            const uint64_t gc = c0 + k / th, f = k % th;
            put(o, hw, (gc / hw) * ths + f * hw + gc % hw, a.col[b + k], a.val[b + k], x, a.ncols);
            if (f == 0) o.sec[gc] = r;
        }
    } else if constexpr (KIND == 2) {
        for (uint32_t k = j; k < l; k += GROUP) {
            const uint32_t p = b + k, bco = p % g.bs;
            put(o, hw, p, a.col[p], a.val[p], x, a.ncols);
            if (k == 0 || bco == 0) o.sec[p] = (r << g.log) | min(l - k - 1, g.bs - 1 - bco);
        }
    } else {
        for (uint32_t k = j; k < l; k += GROUP) {
            const uint64_t p = uint64_t(b) + k, t = p / ths, f = p % th, c = (p - t * ths) / th;
            const uint64_t q = t * ths + f * hw + c;
            put(o, hw, q, a.col[p], a.val[p], x, a.ncols);
            if (k == 0 || f == 0) o.sec[q] = (r << g.log) | min(l - k - 1, th - 1 - uint32_t(f));
        }
    }
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t build_check(const DevCsr &a, uint32_t *flag, hipStream_t s) {
    const uint64_t threads = std::max<uint64_t>(1, std::max<uint64_t>(a.nrows, a.nnz));
    check_kernel<<<unsigned((threads + 255) / 256), 256, 0, s>>>(a, flag);
    return hipGetLastError();
}

hipError_t build_count(const DevCsr &a, uint32_t div, uint64_t *cnt, hipStream_t s) {
    if (a.nrows) count_kernel<<<(a.nrows + 255) / 256, 256, 0, s>>>(a, div, cnt);
    return hipGetLastError();
}

hipError_t build_scan(const uint64_t *in, uint64_t *out, size_t n, void *tmp, size_t &tmp_bytes, hipStream_t s) {
    return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, uint64_t(0), n, rocprim::plus<uint64_t>(), s);
}

hipError_t build_fill(const DevCsr &a, int kind, const uint64_t *start, const BuildGeo &g, const double *x,
                      const BuildOut &o, hipStream_t s) {
    if (!a.nrows) return hipSuccess;
    const uint64_t threads = uint64_t(a.nrows) * GROUP;
    const unsigned grid = unsigned((threads + 255) / 256);
    switch (kind) {
        case 0: fill_kernel<0><<<grid, 256, 0, s>>>(a, start, g, x, o); break;
        case 1: fill_kernel<1><<<grid, 256, 0, s>>>(a, start, g, x, o); break;
        case 2: fill_kernel<2><<<grid, 256, 0, s>>>(a, start, g, x, o); break;
        default: fill_kernel<3><<<grid, 256, 0, s>>>(a, start, g, x, o); break;
    }
    return hipGetLastError();
}

}  // namespace launch
}  // namespace axspmv
