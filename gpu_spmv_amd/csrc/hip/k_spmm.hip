// SpMM: Y = A X with k right-hand sides stored interleaved (X[c*k + j],
// Y[r*k + j]).  Works on any layout whose rows are contiguous runs of slots
// (CSR, the AXT UNC th == 1 planes, AXC): row r's entries sit at plane slots
// first[r] .. first[r] + rl - 1.
//
// A group of kt lanes (the k-tile: 8, 16, 32 or 64) owns one row; lane j of
// the group accumulates right-hand side j.  All lanes of the group read the
// same value and column index (one broadcast load each) and then kt
// neighbouring entries of X -- so every value/index load feeds kt FMAs and the
// X reads of a group are one contiguous kt*8-byte run.  Rows longer than
// `long_slots` are listed on the host and taken by a whole wave, whose groups
// split the row's slots and combine with cross-group DPP/swizzle sums.
#include "launch.hpp"
#include "wave64.hpp"

namespace {
using namespace axspmv::dev;

template <class VT, int KT>
__global__ void __launch_bounds__(256) spmm_rows_kernel(uint32_t nrows, const uint32_t *__restrict__ first,
                                                        const uint32_t *__restrict__ rowptr, const VT *__restrict__ val,
                                                        const uint32_t *__restrict__ col, uint32_t k, uint32_t long_slots,
                                                        const double *__restrict__ X, double *__restrict__ Y) {
    const size_t gl = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gl / KT >= nrows) return;
    const uint32_t r = uint32_t(gl / KT), j = uint32_t(gl % KT);
    const uint32_t kc = blockIdx.y * KT + j;  // right-hand side of this lane
    const uint32_t len = rowptr[r + 1] - rowptr[r];
    if (len > long_slots) return;
    const uint32_t q0 = first[r];
    const bool on = kc < k;
    double acc = 0.0;
    uint32_t t = 0;
    for (; t + 4 <= len; t += 4) {
        const uint32_t q = q0 + t;
        const VT v0 = val[q], v1 = val[q + 1], v2 = val[q + 2], v3 = val[q + 3];
        const uint32_t c0 = col[q], c1 = col[q + 1], c2 = col[q + 2], c3 = col[q + 3];
        if (on) {
            const double x0 = X[size_t(c0) * k + kc], x1 = X[size_t(c1) * k + kc];
            const double x2 = X[size_t(c2) * k + kc], x3 = X[size_t(c3) * k + kc];
            acc = fma(double(v0), x0, acc);
            acc = fma(double(v1), x1, acc);
            acc = fma(double(v2), x2, acc);
            acc = fma(double(v3), x3, acc);
        }
    }
    for (; t < len; ++t) {
        const uint32_t q = q0 + t;
        if (on) acc = fma(double(val[q]), X[size_t(col[q]) * k + kc], acc);
    }
    if (on) Y[size_t(r) * k + kc] = acc;
}

// Sum over the 64/KT groups of a wave (lanes with equal lane % KT).
template <int KT>
__device__ __forceinline__ double cross_group_sum(double v) {
    if constexpr (KT <= 1) v += dpp_f64<DPP_QUAD_1032>(0.0, v);
    if constexpr (KT <= 2) v += dpp_f64<DPP_QUAD_2301>(0.0, v);
    if constexpr (KT <= 4) v += __shfl_xor(v, 4);
    if constexpr (KT <= 8) v += __shfl_xor(v, 8);
    if constexpr (KT <= 16) v += __shfl_xor(v, 16);
    if constexpr (KT <= 32) v += __shfl_xor(v, 32);
    return v;
}

template <class VT, int KT>
__global__ void __launch_bounds__(256) spmm_long_kernel(const uint32_t *__restrict__ rows, uint32_t nlong,
                                                        const uint32_t *__restrict__ first,
                                                        const uint32_t *__restrict__ rowptr, const VT *__restrict__ val,
                                                        const uint32_t *__restrict__ col, uint32_t k,
                                                        const double *__restrict__ X, double *__restrict__ Y) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= nlong) return;
    constexpr uint32_t G = 64 / KT;
    const uint32_t r = rows[wave], g = lane / KT, j = lane % KT;
    const uint32_t kc = blockIdx.y * KT + j;
    const bool on = kc < k;
    const uint32_t q0 = first[r], len = rowptr[r + 1] - rowptr[r];
    double acc = 0.0;
    for (uint32_t t = g; t < len; t += G) {
        const uint32_t q = q0 + t;
        if (on) acc = fma(double(val[q]), X[size_t(col[q]) * k + kc], acc);
    }
    acc = cross_group_sum<KT>(acc);
    if (g == 0 && on) Y[size_t(r) * k + kc] = acc;
}

template <class VT, int KT>
hipError_t run_kt(const axspmv::launch::SpmmArgs &a, uint32_t k, const double *X, double *Y, hipStream_t s) {
    const VT *val = static_cast<const VT *>(a.val);
    const uint32_t ktiles = (k + KT - 1) / KT;
    const uint64_t th = uint64_t(a.nrows) * KT;
    if (a.nrows) {
        dim3 grid(unsigned((th + 255) / 256), ktiles);
        spmm_rows_kernel<VT, KT><<<grid, 256, 0, s>>>(a.nrows, a.first, a.rowptr, val, a.col, k, a.long_slots, X, Y);
    }
    if (a.nlong) {
        dim3 grid((a.nlong + 3) / 4, ktiles);
        spmm_long_kernel<VT, KT><<<grid, 256, 0, s>>>(a.longrows, a.nlong, a.first, a.rowptr, val, a.col, k, X, Y);
    }
    return hipGetLastError();
}

template <class VT>
hipError_t run_vt(const axspmv::launch::SpmmArgs &a, uint32_t k, uint32_t kt, const double *X, double *Y, hipStream_t s) {
    switch (kt) {
        case 8: return run_kt<VT, 8>(a, k, X, Y, s);
        case 16: return run_kt<VT, 16>(a, k, X, Y, s);
        case 32: return run_kt<VT, 32>(a, k, X, Y, s);
        case 64: return run_kt<VT, 64>(a, k, X, Y, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t spmm(const SpmmArgs &a, uint32_t k, uint32_t kt, const double *X, double *Y, hipStream_t s) {
    return a.vtype == 1 ? run_vt<float>(a, k, kt, X, Y, s) : run_vt<double>(a, k, kt, X, Y, s);
}

}  // namespace launch
}  // namespace axspmv
