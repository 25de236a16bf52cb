// CSR, one lane per row: the reference's gcsr (cudaSpmv.cu:531-543) on
// wave64.  y[r] = sum of val*x[col] over the row, left to right from 0.
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
