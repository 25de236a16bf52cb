// The opaque axspmv_plan: device buffers and launch parameters of one matrix
// in one format.
#pragma once

#include <hip/hip_runtime_api.h>

#include <functional>
#include <vector>

#include "../hip/launch.hpp"
#include "common.hpp"

struct axspmv_plan {
    enum Kind { NONE, CSR_NAIVE, AXT_UNC_H1, AXT_UNC_H, AXT_COM, K1, AXC, RCT, AXT_UNC_H1_AX, CSR_VECTOR, AXT_UNC_H1_SLICED, REORDERED };
    Kind kind = NONE;
    int device = 0;
    axspmv_opts o{};
    uint32_t nrows = 0, ncols = 0, nnz = 0;
    char name[64] = "";
    uint64_t slots = 0, device_bytes = 0, stream_bytes = 0, alg_bytes = 0;
    double beta = 0.0;
    uint32_t launches = 1, nblocks = 0, nempty = 0, staged_blocks = 0;
    uint32_t thw = 0, th = 0;
    uint32_t lanes = 0;  // CSR vector: lanes per row

    // CSR arrays
    void *d_val = nullptr;
    uint32_t *d_col = nullptr, *d_row = nullptr;
    // planes + metadata of the tiled formats
    void *d_vplane = nullptr;
    uint32_t *d_cplane = nullptr, *d_sec = nullptr, *d_empty = nullptr;
    axspmv::launch::Window *d_win = nullptr;  // per wave block, when x staging pays
    axspmv::launch::Block *d_blocks = nullptr;
    // K1 work items / AXC long rows
    axspmv::launch::K1Args k1{};
    axspmv::launch::AxcArgs axc{};
    axspmv::launch::ComArgs com{};
    axspmv::launch::RctArgs rct{};
    size_t lds_max = 0;             // LDS bytes per workgroup (RCT SpMM k-tile limit)
    // RCT: keep_pos -- record the CSR position of every entry in rct_pos
    // (set for the transposed plan of a plan with AXSPMV_PLAN_TRANSPOSE; the
    // forward plan then holds the positions of the transposed plan's entries
    // in its own CSR, the order in which v_bar comes out); lds_per_row --
    // LDS bytes a tile row needs (16 for the transposed plan: the x_bar sums
    // and the x values of the fused value gradient).
    bool keep_pos = false;
    uint32_t lds_per_row = 8;
    std::vector<uint32_t> rct_pos;
    // Row-contiguous layouts (CSR, AXT UNC th == 1, AXC): CSR row pointer and
    // first plane slot of each row, for SpMM and the value gradient.
    uint32_t *d_rowptr = nullptr, *d_first = nullptr;
    axspmv::launch::SpmmArgs spmm{};
    bool rowwise = false;
    // AXT UNC th == 1: blocks of <= 64 units for the fused adjoint pass
    axspmv::launch::Block *d_fblocks = nullptr;
    uint32_t nfblocks = 0, nunits = 0;
    uint32_t seg_nt = 0;  // AXT UNC, CSR vector: stream read with non-temporal loads
    // AX layout (AXSPMV_PLAN_AX_LAYOUT): interleaved ax array and the column
    // of every value slot (0xffffffff for padding)
    double *d_ax = nullptr;
    uint32_t *d_axmap = nullptr;
    // AXSPMV_TIME_KERNEL: event the path records after its first kernel (the
    // y clear of RCT and AXT plans in column slices)
    hipEvent_t mid = nullptr;
    // A^T in the same format (AXSPMV_PLAN_TRANSPOSE)
    axspmv_plan *T = nullptr;
    // RCT with AXSPMV_PLAN_DETERMINISTIC: A^T also as a deterministic AXT
    // plan in column slices, for axspmv_adjoint
    axspmv_plan *Tdet = nullptr;
    // AXT in column slices: S, and with AXSPMV_PLAN_DETERMINISTIC the S x
    // nrows slice partials
    uint32_t slices = 1;
    double *d_part = nullptr;
    // AXSPMV_PLAN_REORDER (kind REORDERED): the plan of B = Q A Q^T in
    // `inner`; the permutation into B's order (v~[k] = v[perm[k]]) and back
    // (v[i] = v~[iperm[i]]), each as two launch::permute_chunks passes
    // through a staging vector (pass 1: w1, b1; pass 2: b2; chunks of L);
    // nscratch vectors of nrows doubles (3, or 6 with a transposed plan; the
    // last one stages); bpos[e] = position in A of entry e of B's CSR
    axspmv_plan *inner = nullptr;
    struct Perm {
        uint16_t *b1 = nullptr, *b2 = nullptr;
        uint32_t *w1 = nullptr;
        uint32_t L = 0;
    } to_b, from_b;
    uint32_t *d_perm = nullptr, *d_iperm = nullptr;  // SpMM: rows of X into B's order, rows of Y back
    // SpMM operands in B's order (2 x nrows x k doubles, grown by
    // plan_op_prepare before a call needs them)
    double *d_mm = nullptr;
    uint64_t mm_cap = 0;  // doubles
    double *d_scr[6] = {};
    uint32_t nscratch = 0;
    std::vector<uint32_t> bpos;

    std::vector<void *> bufs;
    void *dalloc(size_t bytes);
    int upload(void *dst, const void *src, size_t bytes, hipStream_t s);
    ~axspmv_plan();
};

namespace axspmv {
using Plan = axspmv_plan;
int hip_fail(hipError_t e, const char *what);
hipError_t plan_launch(axspmv_plan *p, const double *x, double *y, hipStream_t s);
// One operation of enum axspmv_op; plan_op_check validates it first.
hipError_t plan_op(axspmv_plan *p, uint32_t op, const axspmv_operands &o, hipStream_t s);
int plan_op_check(const axspmv_plan *p, uint32_t op, const axspmv_operands &o);
// Device memory an operation needs beyond what the plan holds (a reordered
// plan's SpMM operands for this k); call after plan_op_check and before any
// launch or capture of the operation.
int plan_op_prepare(axspmv_plan *p, uint32_t op, const axspmv_operands &o);
int plan_add_transpose(axspmv_plan *p, const axspmv_csr *a, hipStream_t s);
// Times launches of `fn` (see axspmv_time_op for the modes).
int time_launches(const std::function<hipError_t(hipStream_t)> &fn, uint32_t iters, uint32_t warmup, uint32_t mode,
                  uint32_t graph_batch, void *stream, axspmv_timing *out);
// Captures `repeats` calls of fn into an executable graph.
int capture_graph(const std::function<hipError_t(hipStream_t)> &fn, uint32_t repeats, axspmv_graph **out);
std::vector<launch::Block> row_blocks(const std::vector<uint32_t> &units_per_row, uint32_t budget);
// Column window (smallest to largest column) of every block of
// row_blocks(units_per_row, ...): the blocks hold whole consecutive rows.
std::vector<launch::Window> block_windows(const axspmv_csr *a, const std::vector<uint32_t> &units_per_row,
                                          const std::vector<launch::Block> &blocks);
// True when at least half of the windows fit STAGE_COLS.
bool worth_staging(const std::vector<launch::Window> &w);
}  // namespace axspmv
