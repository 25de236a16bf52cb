// Host-side launch interface of the HIP kernels (implemented in the .hip
// files).  Only plain pointers and sizes cross this boundary.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace axspmv {
namespace launch {

// A wave block: units [beg, end) of one wave (row-aligned for UNC / CSR).
struct Block {
    uint32_t beg, end;
};

// Row-aligned wave-block kernels (AXT UNC th==1 / th>1, CSR stream).
struct SegArgs {
    const void *vplane;       // values (double or float by vtype)
    const uint32_t *cplane;   // column of every slot
    const uint32_t *rowsrc;   // sec[] (UNC) or row pointers (CSR stream)
    const Block *blocks;      // nblocks wave blocks
    uint32_t nblocks;
    const uint32_t *empty;    // rows without entries (y = 0)
    uint32_t nempty;
    uint32_t thw, th;         // AXT tile half width / height
    uint32_t vtype;           // 0 = f64 values, 1 = f32 values
    uint32_t block;           // threads per workgroup
};
hipError_t axt_unc_h1(const SegArgs &a, const double *x, double *y, hipStream_t s);
hipError_t axt_unc_h(const SegArgs &a, const double *x, double *y, hipStream_t s);

hipError_t csr_naive(uint32_t nrows, const void *val, uint32_t vtype, const uint32_t *col, const uint32_t *row,
                     const double *x, double *y, uint32_t block, hipStream_t s);

hipError_t copy_f64(const double *src, double *dst, size_t n, hipStream_t s);
hipError_t fill_f64(double *dst, size_t n, double v, hipStream_t s);

}  // namespace launch
}  // namespace axspmv
