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

}  // namespace

namespace axspmv {
namespace launch {

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
