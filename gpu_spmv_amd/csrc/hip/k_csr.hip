// CSR kernels.  csr_naive: one lane per row, the reference's gcsr
// (cudaSpmv.cu:531-543) on wave64, y[r] = sum of val*x[col] left to right.
// csr_vector: a group of G lanes per row (G divides 64), strided over the
// row, then a DPP group sum; rows longer than long_len entries are listed on
// the host and done one wave per row by csr_long.
#include "launch.hpp"
#include "wave64.hpp"

namespace {

template <class VT>
__global__ void __launch_bounds__(256) csr_naive_kernel(uint32_t nrows, const VT *__restrict__ val,
                                                        const uint32_t *__restrict__ col,
                                                        const uint32_t *__restrict__ row,
                                                        const double *__restrict__ x, double *__restrict__ y) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double acc = 0.0;
    const uint32_t e = row[r + 1];
    for (uint32_t j = row[r]; j < e; ++j) acc += double(val[j]) * x[col[j]];
    y[r] = acc;
}

// NT: val and col are streamed with non-temporal loads.
template <class VT, int G, bool NT>
__global__ void __launch_bounds__(256) csr_vector_kernel(uint32_t nrows, const VT *__restrict__ val,
                                                         const uint32_t *__restrict__ col,
                                                         const uint32_t *__restrict__ row, uint32_t long_len,
                                                         const double *__restrict__ x, double *__restrict__ y) {
    const size_t gl = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gl / G >= nrows) return;  // whole groups leave together (G divides 64)
    const uint32_t r = uint32_t(gl / G), g = uint32_t(gl % G);
    const uint32_t b = row[r], e = row[r + 1];
    if (e - b > long_len) return;  // csr_long's row
    double acc = 0.0;
    for (uint32_t j = b + g; j < e; j += G)
        acc = fma(double(axspmv::dev::stream_load<NT>(val + j)), x[axspmv::dev::stream_load<NT>(col + j)], acc);
    acc = axspmv::dev::group_sum_f64<G>(acc);
    if (g == 0) y[r] = acc;
}

template <class VT, bool NT>
__global__ void __launch_bounds__(256) csr_long_kernel(const uint32_t *__restrict__ rows, uint32_t nlong,
                                                       const VT *__restrict__ val, const uint32_t *__restrict__ col,
                                                       const uint32_t *__restrict__ row, const double *__restrict__ x,
                                                       double *__restrict__ y) {
    const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= nlong) return;
    const uint32_t r = rows[w];
    double acc = 0.0;
    for (uint32_t j = row[r] + lane; j < row[r + 1]; j += 64)
        acc = fma(double(axspmv::dev::stream_load<NT>(val + j)), x[axspmv::dev::stream_load<NT>(col + j)], acc);
    acc = axspmv::dev::group_sum_f64<64>(acc);
    if (lane == 0) y[r] = acc;
}

template <class VT, bool NT>
hipError_t csr_vector_vt(uint32_t nrows, const VT *val, const uint32_t *col, const uint32_t *row, uint32_t lanes,
                         uint32_t long_len, const uint32_t *longrows, uint32_t nlong, const double *x, double *y,
                         hipStream_t s) {
    const unsigned grid = unsigned((uint64_t(nrows) * lanes + 255) / 256);
    switch (lanes) {
#define AX_G_CASE(G) \
    case G: csr_vector_kernel<VT, G, NT><<<grid, 256, 0, s>>>(nrows, val, col, row, long_len, x, y); break;
        AX_G_CASE(1)
        AX_G_CASE(2)
        AX_G_CASE(4)
        AX_G_CASE(8)
        AX_G_CASE(16)
        AX_G_CASE(32)
        AX_G_CASE(64)
#undef AX_G_CASE
        default: return hipErrorInvalidValue;
    }
    if (nlong) csr_long_kernel<VT, NT><<<(nlong + 3) / 4, 256, 0, s>>>(longrows, nlong, val, col, row, x, y);
    return hipGetLastError();
}

// v_bar[k] = y_bar[row(k)] * x[col(k)], one lane per row.
__global__ void __launch_bounds__(256) csr_vbar_kernel(uint32_t nrows, const uint32_t *__restrict__ col,
                                                       const uint32_t *__restrict__ row, const double *__restrict__ x,
                                                       const double *__restrict__ ybar, double *__restrict__ vbar) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const double s = ybar[r];
    for (uint32_t j = row[r]; j < row[r + 1]; ++j) vbar[j] = s * x[col[j]];
}

// y = A x, r = y - b, v_bar = r[row] * x[col], one lane per row.
template <class VT>
__global__ void __launch_bounds__(256) csr_fused_kernel(uint32_t nrows, const VT *__restrict__ val,
                                                        const uint32_t *__restrict__ col,
                                                        const uint32_t *__restrict__ row, const double *__restrict__ x,
                                                        const double *__restrict__ b, double *__restrict__ y,
                                                        double *__restrict__ res, double *__restrict__ vbar) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double acc = 0.0;
    const uint32_t e = row[r + 1];
    for (uint32_t j = row[r]; j < e; ++j) acc += double(val[j]) * x[col[j]];
    const double rv = acc - b[r];
    y[r] = acc;
    res[r] = rv;
    for (uint32_t j = row[r]; j < e; ++j) vbar[j] = rv * x[col[j]];
}

__global__ void __launch_bounds__(256) residual_kernel(uint32_t n, const double *__restrict__ y,
                                                       const double *__restrict__ b, double *__restrict__ r) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = y[i] - b[i];
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t csr_vbar(uint32_t nrows, const uint32_t *col, const uint32_t *row, const double *x, const double *ybar,
                    double *vbar, hipStream_t s) {
    if (nrows) csr_vbar_kernel<<<(nrows + 255) / 256, 256, 0, s>>>(nrows, col, row, x, ybar, vbar);
    return hipGetLastError();
}

hipError_t csr_fused(uint32_t nrows, const void *val, uint32_t vtype, const uint32_t *col, const uint32_t *row,
                     const double *x, const double *b, double *y, double *res, double *vbar, hipStream_t s) {
    if (!nrows) return hipSuccess;
    const unsigned grid = (nrows + 255) / 256;
    if (vtype == 1)
        csr_fused_kernel<float><<<grid, 256, 0, s>>>(nrows, static_cast<const float *>(val), col, row, x, b, y, res, vbar);
    else
        csr_fused_kernel<double><<<grid, 256, 0, s>>>(nrows, static_cast<const double *>(val), col, row, x, b, y, res, vbar);
    return hipGetLastError();
}

hipError_t residual(uint32_t n, const double *y, const double *b, double *r, hipStream_t s) {
    if (n) residual_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, y, b, r);
    return hipGetLastError();
}

hipError_t csr_vector(uint32_t nrows, const void *val, uint32_t vtype, const uint32_t *col, const uint32_t *row,
                      uint32_t lanes, uint32_t long_len, const uint32_t *longrows, uint32_t nlong, const double *x,
                      double *y, hipStream_t s, uint32_t nt) {
    if (!nrows) return hipSuccess;
    const float *vf = static_cast<const float *>(val);
    const double *vd = static_cast<const double *>(val);
    if (nt)
        return vtype == 1 ? csr_vector_vt<float, true>(nrows, vf, col, row, lanes, long_len, longrows, nlong, x, y, s)
                          : csr_vector_vt<double, true>(nrows, vd, col, row, lanes, long_len, longrows, nlong, x, y, s);
    return vtype == 1 ? csr_vector_vt<float, false>(nrows, vf, col, row, lanes, long_len, longrows, nlong, x, y, s)
                      : csr_vector_vt<double, false>(nrows, vd, col, row, lanes, long_len, longrows, nlong, x, y, s);
}

hipError_t csr_naive(uint32_t nrows, const void *val, uint32_t vtype, const uint32_t *col, const uint32_t *row,
                     const double *x, double *y, uint32_t block, hipStream_t s) {
    if (!nrows) return hipSuccess;
    if (block == 0 || block > 256 || block % 64) block = 256;
    const unsigned grid = (nrows + block - 1) / block;
    if (vtype == 1)
        csr_naive_kernel<float><<<grid, block, 0, s>>>(nrows, static_cast<const float *>(val), col, row, x, y);
    else
        csr_naive_kernel<double><<<grid, block, 0, s>>>(nrows, static_cast<const double *>(val), col, row, x, y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace axspmv
