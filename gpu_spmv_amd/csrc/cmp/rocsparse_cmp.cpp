// rocSPARSE comparator for the benchmark -- not part of libaxspmv.
//
// The reference benchmarks the vendor library beside its own kernels
// (test_gcucsr, cudaSpmv.cu:629-704: cusparseCsrmvEx, NUM_ITE launches, one
// event pair per launch).  This is the MI355X counterpart: rocSPARSE's CSR
// SpMV (the v2 generic API) on the same device CSR arrays, built into its own
// librocsparse_cmp.so that only bench.py and the tests load.  No library
// path dispatches to it.
//
//   int cmp_rocsparse_spmv(nrows, ncols, nnz, d_row, d_col, d_val, d_x, d_y,
//                          alg, iters, &analysis_ms, &mean_ms, &api)
//     y = A x with algorithm alg (0 adaptive, 1 rowsplit, 2 LRB, 3
//     nnzsplit): the analysis stage once (timed on the host clock, it
//     blocks), two warm-up computes, then iters computes between one event
//     pair (mean per SpMV).  The v2 generic API first, the original one if
//     v2 refuses (api = 2 or 1).  0 or a negative code; cmp_last_error()
//     says why.
#include <hip/hip_runtime_api.h>
#include <rocsparse/rocsparse.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <string>

namespace {

thread_local char g_err[256] = "";

int fail(const char *what, int code) {
    std::snprintf(g_err, sizeof g_err, "%s failed (%d)", what, code);
    return -1;
}

}  // namespace

#define CMP_SP(expr)                                                  \
    do {                                                              \
        const rocsparse_status s_ = (expr);                           \
        if (s_ != rocsparse_status_success) { rc = fail(#expr, s_); goto done; } \
    } while (0)
#define CMP_HIP(expr)                                                 \
    do {                                                              \
        const hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) { rc = fail(#expr, e_); goto done; }    \
    } while (0)

extern "C" const char *cmp_last_error(void) { return g_err; }

namespace {

// The v2 generic API: analysis (its own buffer, sized first), then compute
// (its buffer sized after the analysis).  api_out = 2.
int run_v2(rocsparse_handle h, rocsparse_spmat_descr A, rocsparse_dnvec_descr X, rocsparse_dnvec_descr Y,
           rocsparse_spmv_alg alg, uint32_t iters, double *analysis_ms, double *mean_ms) {
    int rc = 0;
    rocsparse_spmv_descr D = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void *abuf = nullptr, *cbuf = nullptr;
    size_t abytes = 0, cbytes = 0;
    const double one = 1.0, zero = 0.0;
    const rocsparse_operation op = rocsparse_operation_none;
    const rocsparse_datatype f64 = rocsparse_datatype_f64_r;
    float ms = 0.0f;
    CMP_SP(rocsparse_create_spmv_descr(&D));
    CMP_SP(rocsparse_spmv_set_input(h, D, rocsparse_spmv_input_alg, &alg, sizeof alg, nullptr));
    CMP_SP(rocsparse_spmv_set_input(h, D, rocsparse_spmv_input_operation, &op, sizeof op, nullptr));
    CMP_SP(rocsparse_spmv_set_input(h, D, rocsparse_spmv_input_scalar_datatype, &f64, sizeof f64, nullptr));
    CMP_SP(rocsparse_spmv_set_input(h, D, rocsparse_spmv_input_compute_datatype, &f64, sizeof f64, nullptr));
    CMP_SP(rocsparse_v2_spmv_buffer_size(h, D, A, X, Y, rocsparse_v2_spmv_stage_analysis, &abytes, nullptr));
    if (abytes) CMP_HIP(hipMalloc(&abuf, abytes));
    {
        const auto t0 = std::chrono::steady_clock::now();
        CMP_SP(rocsparse_v2_spmv(h, D, &one, A, X, &zero, Y, rocsparse_v2_spmv_stage_analysis, abytes, abuf, nullptr));
        CMP_HIP(hipDeviceSynchronize());
        *analysis_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    CMP_SP(rocsparse_v2_spmv_buffer_size(h, D, A, X, Y, rocsparse_v2_spmv_stage_compute, &cbytes, nullptr));
    if (cbytes) CMP_HIP(hipMalloc(&cbuf, cbytes));
    for (int w = 0; w < 2; ++w)
        CMP_SP(rocsparse_v2_spmv(h, D, &one, A, X, &zero, Y, rocsparse_v2_spmv_stage_compute, cbytes, cbuf, nullptr));
    CMP_HIP(hipEventCreate(&e0));
    CMP_HIP(hipEventCreate(&e1));
    CMP_HIP(hipEventRecord(e0, nullptr));
    for (uint32_t i = 0; i < iters; ++i)
        CMP_SP(rocsparse_v2_spmv(h, D, &one, A, X, &zero, Y, rocsparse_v2_spmv_stage_compute, cbytes, cbuf, nullptr));
    CMP_HIP(hipEventRecord(e1, nullptr));
    CMP_HIP(hipEventSynchronize(e1));
    CMP_HIP(hipEventElapsedTime(&ms, e0, e1));
    *mean_ms = double(ms) / iters;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (abuf) (void)hipFree(abuf);
    if (cbuf) (void)hipFree(cbuf);
    if (D) (void)rocsparse_destroy_spmv_descr(D);
    return rc;
}

// The original generic API (buffer size, preprocess, compute on one
// buffer), still exported and deprecated in favour of v2.  api_out = 1.
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdeprecated-declarations"
int run_v1(rocsparse_handle h, rocsparse_spmat_descr A, rocsparse_dnvec_descr X, rocsparse_dnvec_descr Y,
           rocsparse_spmv_alg alg, uint32_t iters, double *analysis_ms, double *mean_ms) {
    int rc = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void *buf = nullptr;
    size_t bytes = 0;
    const double one = 1.0, zero = 0.0;
    const rocsparse_operation op = rocsparse_operation_none;
    const rocsparse_datatype f64 = rocsparse_datatype_f64_r;
    float ms = 0.0f;
    CMP_SP(rocsparse_spmv(h, op, &one, A, X, &zero, Y, f64, alg, rocsparse_spmv_stage_buffer_size, &bytes, nullptr));
    CMP_HIP(hipMalloc(&buf, bytes ? bytes : 256));
    {
        const auto t0 = std::chrono::steady_clock::now();
        CMP_SP(rocsparse_spmv(h, op, &one, A, X, &zero, Y, f64, alg, rocsparse_spmv_stage_preprocess, &bytes, buf));
        CMP_HIP(hipDeviceSynchronize());
        *analysis_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    for (int w = 0; w < 2; ++w)
        CMP_SP(rocsparse_spmv(h, op, &one, A, X, &zero, Y, f64, alg, rocsparse_spmv_stage_compute, &bytes, buf));
    CMP_HIP(hipEventCreate(&e0));
    CMP_HIP(hipEventCreate(&e1));
    CMP_HIP(hipEventRecord(e0, nullptr));
    for (uint32_t i = 0; i < iters; ++i)
        CMP_SP(rocsparse_spmv(h, op, &one, A, X, &zero, Y, f64, alg, rocsparse_spmv_stage_compute, &bytes, buf));
    CMP_HIP(hipEventRecord(e1, nullptr));
    CMP_HIP(hipEventSynchronize(e1));
    CMP_HIP(hipEventElapsedTime(&ms, e0, e1));
    *mean_ms = double(ms) / iters;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (buf) (void)hipFree(buf);
    return rc;
}
#pragma GCC diagnostic pop

}  // namespace

extern "C" int cmp_rocsparse_spmv(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *d_row,
                                  const uint32_t *d_col, const double *d_val, const double *d_x, double *d_y,
                                  int alg, uint32_t iters, double *analysis_ms, double *mean_ms, int *api_out) {
    static const rocsparse_spmv_alg algs[] = {rocsparse_spmv_alg_csr_adaptive, rocsparse_spmv_alg_csr_rowsplit,
                                              rocsparse_spmv_alg_csr_lrb, rocsparse_spmv_alg_csr_nnzsplit};
    if (alg < 0 || alg > 3 || !analysis_ms || !mean_ms || !api_out || iters == 0) {
        std::snprintf(g_err, sizeof g_err, "bad argument");
        return -2;
    }
    int rc = 0;
    rocsparse_handle h = nullptr;
    rocsparse_spmat_descr A = nullptr;
    rocsparse_dnvec_descr X = nullptr, Y = nullptr;
    const rocsparse_datatype f64 = rocsparse_datatype_f64_r;
    // an error left pending by earlier HIP calls of the process is reported
    // (and cleared) rather than mistaken for rocSPARSE's own
    const hipError_t stale = hipGetLastError();
    if (rocsparse_create_handle(&h) != rocsparse_status_success) {
        std::snprintf(g_err, sizeof g_err, "rocsparse_create_handle failed (pending HIP error before it: %d %s)",
                      int(stale), hipGetErrorName(stale));
        return -1;
    }
    CMP_SP(rocsparse_create_csr_descr(&A, nrows, ncols, nnz, const_cast<uint32_t *>(d_row),
                                      const_cast<uint32_t *>(d_col), const_cast<double *>(d_val),
                                      rocsparse_indextype_i32, rocsparse_indextype_i32, rocsparse_index_base_zero,
                                      f64));
    CMP_SP(rocsparse_create_dnvec_descr(&X, ncols, const_cast<double *>(d_x), f64));
    CMP_SP(rocsparse_create_dnvec_descr(&Y, nrows, d_y, f64));
    *api_out = 2;
    rc = run_v2(h, A, X, Y, algs[alg], iters, analysis_ms, mean_ms);
    if (rc) {
        const std::string v2 = g_err;
        *api_out = 1;
        rc = run_v1(h, A, X, Y, algs[alg], iters, analysis_ms, mean_ms);
        if (rc) {
            const std::string v1 = g_err;
            std::snprintf(g_err, sizeof g_err, "v2: %s; v1: %s", v2.c_str(), v1.c_str());
        }
    }
done:
    if (X) (void)rocsparse_destroy_dnvec_descr(X);
    if (Y) (void)rocsparse_destroy_dnvec_descr(Y);
    if (A) (void)rocsparse_destroy_spmat_descr(A);
    if (h) (void)rocsparse_destroy_handle(h);
    return rc;
}
