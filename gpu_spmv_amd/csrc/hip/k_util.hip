// Utility kernels: a plain copy (the bandwidth reference used to calibrate
// the HBM byte counters), fills, the x-half refresh of the interleaved ax
// layouts, and a lane-by-lane exercise of the wave64 primitives.
#include "launch.hpp"
#include "wave64.hpp"

namespace {

__global__ void __launch_bounds__(256) copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n2) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += stride) dst[i] = src[i];
}

__global__ void copy_tail(const double *__restrict__ src, double *__restrict__ dst, size_t n) {
    if (threadIdx.x == 0 && (n & 1)) dst[n - 1] = src[n - 1];
}

__global__ void __launch_bounds__(256) fill_kernel(double *__restrict__ dst, size_t n, double v) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = v;
}

// y[0 .. n) = 0, one lane per 16 bytes (VEC: y 16-byte aligned) or per
// element: the clear before the atomic slice adds of the sliced plans (one
// pass over y, no grid-stride loop).  Non-temporal stores: the zeros go to
// memory without being kept in the clearing XCD's L2, and the slice adds
// that follow, from every XCD, run faster for it (RCT on the 2^20-row
// power-law matrix: tile kernel 69.5 -> 64.4 us, clear + tile kernel
// 71.8 -> 66.6 us; the clear itself takes the same 5 us).
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <bool VEC>
__global__ void __launch_bounds__(256) clear_kernel(double *__restrict__ y, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if constexpr (VEC) {
        if (2 * i + 1 < n)
            __builtin_nontemporal_store(f64x2{0.0, 0.0}, reinterpret_cast<f64x2 *>(y) + i);
        else if (2 * i < n)
            __builtin_nontemporal_store(0.0, y + 2 * i);
    } else {
        if (i < n) __builtin_nontemporal_store(0.0, y + i);
    }
}

// y[r] = sum of the S slice partials of row r, in slice order; two rows per
// lane (VEC: part and y 16-byte aligned, nrows even).  Non-temporal loads
// and stores: the partials are read once, and y kept in the L2s would take
// the room the next SpMV's x slices need (RCT on the 2^20-row power-law
// matrix, y stored normally: 61.7 or 78.5 us per SpMV depending on where y
// lies, as with a y cleared by ordinary stores).
// With res (a forward + adjoint pass), res = y - rb as well, stored
// normally: the adjoint reads it next.
template <bool VEC>
__global__ void __launch_bounds__(256) merge_kernel(const double *__restrict__ part, uint32_t S, uint32_t nrows,
                                                    double *__restrict__ y, const double *__restrict__ rb,
                                                    double *__restrict__ res) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (VEC) {
        if (2 * i >= nrows) return;
        const f64x2 *p2 = reinterpret_cast<const f64x2 *>(part) + i;
        f64x2 v = __builtin_nontemporal_load(p2);
        for (uint32_t k = 1; k < S; ++k) v += __builtin_nontemporal_load(p2 + size_t(k) * (nrows / 2));
        __builtin_nontemporal_store(v, reinterpret_cast<f64x2 *>(y) + i);
        if (res) reinterpret_cast<f64x2 *>(res)[i] = v - reinterpret_cast<const f64x2 *>(rb)[i];
    } else {
        if (i >= nrows) return;
        double v = __builtin_nontemporal_load(part + i);
        for (uint32_t k = 1; k < S; ++k) v += __builtin_nontemporal_load(part + size_t(k) * nrows + i);
        __builtin_nontemporal_store(v, y + i);
        if (res) res[i] = v - rb[i];
    }
}

__global__ void __launch_bounds__(256) fill_u32_kernel(uint32_t *__restrict__ dst, size_t n, uint32_t v) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = v;
}

// One chunk of a permutation per workgroup, through LDS: chunk c covers
// positions [c L, c L + len) of src; lds[i] = src[c L + i] (read whole), then
// dst[w ? w[c L + i] : c L + i] = lds[b[c L + i]].  A permutation runs as two
// such launches through a staging vector (the host arranges w and b): the
// first reads a block of the source and writes its values in runs, one run
// per destination chunk; the second reads a destination chunk's values, now
// side by side, and writes the chunk in order.  No lane reads memory at
// random; each lane issues U loads before it waits.
template <bool W>
__global__ void __launch_bounds__(1024) permute_chunk_kernel(const uint32_t *__restrict__ w,
                                                             const uint16_t *__restrict__ b, uint32_t n, uint32_t L,
                                                             const double *__restrict__ src,
                                                             double *__restrict__ dst) {
    constexpr int U = 8;
    extern __shared__ double lds[];
    const uint64_t base = uint64_t(blockIdx.x) * L;
    const uint32_t len = n - base < L ? uint32_t(n - base) : L, nt = blockDim.x;
    for (uint32_t i0 = threadIdx.x; i0 < len; i0 += U * nt) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            if (i < len) v[u] = src[base + i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            if (i < len) lds[i] = v[u];
        }
    }
    __syncthreads();
    for (uint32_t i0 = threadIdx.x; i0 < len; i0 += U * nt) {
        uint32_t j[U], to[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            j[u] = i < len ? b[base + i] : 0;
            if constexpr (W) to[u] = i < len ? w[base + i] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            if (i < len) dst[W ? to[u] : base + i] = lds[j[u]];
        }
    }
}

// The second pass of one permutation (y[c L + i] = t[c L + b2[c L + i]]),
// the residual r = y - bv, and the first pass of another that takes r from
// the same chunks (t2[w1[c L + i]] = r[c L + b1[c L + i]]), one chunk per
// workgroup: y's chunk, then r's, stay in LDS (2 L doubles), so r is not
// read back.
__global__ void __launch_bounds__(1024) back_residual_in_kernel(const uint16_t *__restrict__ b2,
                                                                const uint32_t *__restrict__ w1,
                                                                const uint16_t *__restrict__ b1, uint32_t n,
                                                                uint32_t L, const double *__restrict__ t,
                                                                const double *__restrict__ bv, double *__restrict__ y,
                                                                double *__restrict__ r, double *__restrict__ t2) {
    constexpr int U = 8;
    extern __shared__ double lds[];
    double *lr = lds + L;
    const uint64_t base = uint64_t(blockIdx.x) * L;
    const uint32_t len = n - base < L ? uint32_t(n - base) : L, nt = blockDim.x;
    for (uint32_t i0 = threadIdx.x; i0 < len; i0 += U * nt) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            if (i < len) v[u] = t[base + i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            if (i < len) lds[i] = v[u];
        }
    }
    __syncthreads();
    for (uint32_t i0 = threadIdx.x; i0 < len; i0 += U * nt) {
        uint32_t j[U];
        double bb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            j[u] = i < len ? b2[base + i] : 0;
            bb[u] = i < len ? bv[base + i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            if (i < len) {
                const double v = lds[j[u]], rv = v - bb[u];
                y[base + i] = v;
                r[base + i] = rv;
                lr[i] = rv;
            }
        }
    }
    __syncthreads();
    for (uint32_t i0 = threadIdx.x; i0 < len; i0 += U * nt) {
        uint32_t j[U], to[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            j[u] = i < len ? b1[base + i] : 0;
            to[u] = i < len ? w1[base + i] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * nt;
            if (i < len) t2[to[u]] = lr[j[u]];
        }
    }
}

// dst[i k + j] = src[idx[i] k + j]: rows of k values moved whole (a
// reordered plan's SpMM operands), one lane per value, a row's lanes side by
// side.
__global__ void __launch_bounds__(256) gather_rows_kernel(const uint32_t *__restrict__ idx, uint64_t total,
                                                          uint32_t k, const double *__restrict__ src,
                                                          double *__restrict__ dst) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
        const size_t i = t / k, j = t - i * k;
        dst[t] = src[size_t(idx[i]) * k + j];
    }
}

__global__ void __launch_bounds__(256) fill_rows_kernel(const uint32_t *__restrict__ rows, uint32_t n, double *__restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[rows[i]] = 0.0;
}

// One lane per value slot: map loads and ax stores are contiguous within each
// group of hw slots; x is gathered.
__global__ void __launch_bounds__(256) x_refresh_kernel(const uint32_t *__restrict__ map, size_t n, uint32_t hw_log,
                                                        uint32_t nx, const double *__restrict__ x,
                                                        double *__restrict__ ax) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t hw = size_t(1) << hw_log;
    for (size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += stride) {
        const uint32_t c = map[q];
        ax[((q >> hw_log) << (hw_log + 1)) + hw + (q & (hw - 1))] = c < nx ? x[c] : 0.0;
    }
}

// Reads n2 double2 (a cache flush that leaves no dirty lines); the store
// happens only for a sum no real data produces, so the loads stay.
__global__ void __launch_bounds__(256) touch_kernel(const double2 *__restrict__ p, size_t n2, double *__restrict__ sink) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    double acc = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += stride) {
        const double2 v = p[i];
        acc += v.x + v.y;
    }
    if (acc == -1.0e300) sink[0] = acc;
}

unsigned grid_for(size_t work, unsigned block) {
    size_t g = (work + block - 1) / block;
    const size_t cap = 256 * 64;  // 256 CUs x 64 workgroups: enough to fill the chip
    if (g > cap) g = cap;
    return g ? unsigned(g) : 1u;
}

// One lane per element, whole waves (n a multiple of 64): applies the wave64
// primitive that op names (enum axspmv_wave_op) to v / h and stores the
// lane's results, so each primitive can be checked against its definition.
__global__ void __launch_bounds__(256) wave_selftest_kernel(uint32_t op, const double *__restrict__ v,
                                                           const uint32_t *__restrict__ h, double *__restrict__ ov,
                                                           uint32_t *__restrict__ oh, size_t n) {
    using namespace axspmv::dev;
    // grid stride (a multiple of 64) over whole waves: n is a multiple of 64,
    // so the lanes of a wave stay together and leave together
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        double a = v[i];
        uint32_t b = h[i];
        switch (op) {  // op is uniform: every lane of the wave takes the same case
        case 0: a = seg_scan_f64(a, b); break;
        case 1: b = scan_max_u32(b); break;
        case 2:
            a = shift_up_f64(a, -1.0);
            b = shift_up_u32(b, 0xffffffffu);
            break;
        case 3: b = shift_down_u32(b, 0xffffffffu); break;
        case 4: a = group_sum_f64<1>(a); break;
        case 5: a = group_sum_f64<2>(a); break;
        case 6: a = group_sum_f64<4>(a); break;
        case 7: a = group_sum_f64<8>(a); break;
        case 8: a = group_sum_f64<16>(a); break;
        case 9: a = group_sum_f64<32>(a); break;
        default: a = group_sum_f64<64>(a); break;
        }
        ov[i] = a;
        oh[i] = b;
    }
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t copy_f64(const double *src, double *dst, size_t n, hipStream_t s) {
    const size_t n2 = n / 2;
    if (n2) copy_kernel<<<grid_for(n2, 256), 256, 0, s>>>(reinterpret_cast<const double2 *>(src),
                                                            reinterpret_cast<double2 *>(dst), n2);
    if (n & 1) copy_tail<<<1, 64, 0, s>>>(src, dst, n);
    return hipGetLastError();
}

hipError_t fill_f64(double *dst, size_t n, double v, hipStream_t s) {
    if (n) fill_kernel<<<grid_for(n, 256), 256, 0, s>>>(dst, n, v);
    return hipGetLastError();
}

hipError_t x_refresh(const uint32_t *map, uint64_t n, uint32_t hw_log, uint32_t nx, const double *x, double *ax,
                     hipStream_t s) {
    if (n) x_refresh_kernel<<<grid_for(n, 256), 256, 0, s>>>(map, n, hw_log, nx, x, ax);
    return hipGetLastError();
}

hipError_t wave_selftest(uint32_t op, const double *v, const uint32_t *h, double *ov, uint32_t *oh, size_t n,
                         hipStream_t s) {
    if (n) wave_selftest_kernel<<<grid_for(n, 256), 256, 0, s>>>(op, v, h, ov, oh, n);
    return hipGetLastError();
}

hipError_t touch(const double *p, size_t n, double *sink, hipStream_t s) {
    if (n / 2) touch_kernel<<<grid_for(n / 2, 256), 256, 0, s>>>(reinterpret_cast<const double2 *>(p), n / 2, sink);
    return hipGetLastError();
}

hipError_t clear_f64(double *y, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    if (reinterpret_cast<uintptr_t>(y) % 16 == 0)
        clear_kernel<true><<<unsigned((n / 2 + 256) / 256), 256, 0, s>>>(y, n);
    else
        clear_kernel<false><<<unsigned((n + 255) / 256), 256, 0, s>>>(y, n);
    return hipGetLastError();
}

hipError_t part_merge(const double *part, uint32_t S, uint32_t nrows, double *y, hipStream_t s, const double *rb,
                      double *res) {
    if (!nrows) return hipSuccess;
    auto a16 = [](const void *q) { return reinterpret_cast<uintptr_t>(q) % 16 == 0; };
    if (nrows % 2 == 0 && a16(part) && a16(y) && (!res || (a16(res) && a16(rb))))
        merge_kernel<true><<<(nrows / 2 + 255) / 256, 256, 0, s>>>(part, S, nrows, y, rb, res);
    else
        merge_kernel<false><<<(nrows + 255) / 256, 256, 0, s>>>(part, S, nrows, y, rb, res);
    return hipGetLastError();
}

hipError_t fill_u32(uint32_t *dst, size_t n, uint32_t v, hipStream_t s) {
    if (n) fill_u32_kernel<<<grid_for(n, 256), 256, 0, s>>>(dst, n, v);
    return hipGetLastError();
}

hipError_t back_residual_in(const uint16_t *b2, const uint32_t *w1, const uint16_t *b1, uint32_t n, uint32_t L,
                            const double *t, const double *bv, double *y, double *r, double *t2, hipStream_t s) {
    const unsigned grid = unsigned((uint64_t(n) + L - 1) / L);
    if (n) back_residual_in_kernel<<<grid, 1024, size_t(L) * 16, s>>>(b2, w1, b1, n, L, t, bv, y, r, t2);
    return hipGetLastError();
}

hipError_t permute_chunks(const uint32_t *w, const uint16_t *b, uint32_t n, uint32_t L, const double *src,
                          double *dst, hipStream_t s) {
    const unsigned grid = unsigned((uint64_t(n) + L - 1) / L);
    if (n && w)
        permute_chunk_kernel<true><<<grid, 1024, size_t(L) * 8, s>>>(w, b, n, L, src, dst);
    else if (n)
        permute_chunk_kernel<false><<<grid, 1024, size_t(L) * 8, s>>>(w, b, n, L, src, dst);
    return hipGetLastError();
}

hipError_t gather_rows(const uint32_t *idx, uint32_t n, uint32_t k, const double *src, double *dst, hipStream_t s) {
    const uint64_t total = uint64_t(n) * k;
    if (total) gather_rows_kernel<<<grid_for(total, 256), 256, 0, s>>>(idx, total, k, src, dst);
    return hipGetLastError();
}

hipError_t fill_rows(const uint32_t *rows, uint32_t n, double *y, hipStream_t s) {
    if (n) fill_rows_kernel<<<(n + 255) / 256, 256, 0, s>>>(rows, n, y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace axspmv
