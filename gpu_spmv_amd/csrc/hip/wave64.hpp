// Wave64 cross-lane primitives for gfx950 built on DPP (data-parallel
// primitives: row_shr, row_bcast15/31, wave_shr/shl, mirrors) and readlane.
//
// They replace the 32-lane __shfl_* trees and Hillis-Steele scans of the
// reference (cudaSpmv.cu:1086-1090, 1586-1594, 2061-2086; scan.cu:239-414).
// A 64-bit value moves as two 32-bit DPP movs with the same control.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axspmv {
namespace dev {

constexpr uint32_t NOROW = 0xffffffffu;

// x[c] for a slot of a padded layout: the plans mark padding slots with the
// column 0xffffffff (AXSPMV_PAD_COL), which reads 0.0, so a non-finite x
// entry cannot reach a row through a padding slot (0 * inf = NaN).
__device__ __forceinline__ double xpad(const double *__restrict__ x, uint32_t c) {
    return c == 0xffffffffu ? 0.0 : x[c];
}

// One load of a matrix stream; NT: non-temporal (the plans ask for it where
// the stream passes through the Infinity Cache without staying, so its
// lines need not evict the x lines the gathers reuse).
template <bool NT, class T>
__device__ __forceinline__ T stream_load(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// DPP control words (GFX9 encoding).
enum : int {
    DPP_ROW_SHR1 = 0x111,
    DPP_ROW_SHR2 = 0x112,
    DPP_ROW_SHR4 = 0x114,
    DPP_ROW_SHR8 = 0x118,
    DPP_ROW_BCAST15 = 0x142,
    DPP_ROW_BCAST31 = 0x143,
    DPP_WAVE_SHL1 = 0x130,
    DPP_WAVE_SHR1 = 0x138,
    DPP_ROW_MIRROR = 0x140,
    DPP_ROW_HALF_MIRROR = 0x141,
    DPP_QUAD_1032 = 0xB1,  // quad_perm [1,0,3,2]
    DPP_QUAD_2301 = 0x4E,  // quad_perm [2,3,0,1]
};

// Lanes whose DPP source is out of range or whose row is masked keep `old`.
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWMASK, 0xf, false);
}

template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp_f64(double old, double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, ROWMASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ uint32_t readfirst_u32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// Lane i receives lane i-1's value; lane 0 receives `fill`.
__device__ __forceinline__ uint32_t shift_up_u32(uint32_t v, uint32_t fill) { return dpp_u32<DPP_WAVE_SHR1>(fill, v); }
__device__ __forceinline__ double shift_up_f64(double v, double fill) { return dpp_f64<DPP_WAVE_SHR1>(fill, v); }
// Lane i receives lane i+1's value; lane 63 receives `fill`.
__device__ __forceinline__ uint32_t shift_down_u32(uint32_t v, uint32_t fill) { return dpp_u32<DPP_WAVE_SHL1>(fill, v); }

// Segmented inclusive sum over the wave.  On entry `head` is nonzero on the
// lanes that start a segment; on exit v is the sum from the lane's segment
// head (or lane 0) to the lane, and `head` is nonzero iff a head lies in
// [0, lane].  Hillis-Steele inside 16-lane rows, then row_bcast15/31.
__device__ __forceinline__ double seg_scan_f64(double v, uint32_t &head) {
#define AX_SEG_STEP(CTRL, RM)                          \
    {                                                  \
        const double sv = dpp_f64<CTRL, RM>(0.0, v);   \
        const uint32_t sf = dpp_u32<CTRL, RM>(0u, head); \
        v = head ? v : v + sv;                         \
        head |= sf;                                    \
    }
    AX_SEG_STEP(DPP_ROW_SHR1, 0xf)
    AX_SEG_STEP(DPP_ROW_SHR2, 0xf)
    AX_SEG_STEP(DPP_ROW_SHR4, 0xf)
    AX_SEG_STEP(DPP_ROW_SHR8, 0xf)
    AX_SEG_STEP(DPP_ROW_BCAST15, 0xa)
    AX_SEG_STEP(DPP_ROW_BCAST31, 0xc)
#undef AX_SEG_STEP
    return v;
}

// Inclusive max over the wave of unsigned values (used to propagate the
// current row id forward: rows are nondecreasing in slot order).
__device__ __forceinline__ uint32_t scan_max_u32(uint32_t v) {
    v = max(v, dpp_u32<DPP_ROW_SHR1>(0u, v));
    v = max(v, dpp_u32<DPP_ROW_SHR2>(0u, v));
    v = max(v, dpp_u32<DPP_ROW_SHR4>(0u, v));
    v = max(v, dpp_u32<DPP_ROW_SHR8>(0u, v));
    v = max(v, dpp_u32<DPP_ROW_BCAST15, 0xa>(0u, v));
    v = max(v, dpp_u32<DPP_ROW_BCAST31, 0xc>(0u, v));
    return v;
}

// Sum over aligned groups of G lanes (G in 1,2,4,8,16,32,64); every lane of
// a group ends with the group total.
template <int G>
__device__ __forceinline__ double group_sum_f64(double v) {
    if constexpr (G >= 2) v += dpp_f64<DPP_QUAD_1032>(0.0, v);
    if constexpr (G >= 4) v += dpp_f64<DPP_QUAD_2301>(0.0, v);
    if constexpr (G >= 8) v += dpp_f64<DPP_ROW_HALF_MIRROR>(0.0, v);
    if constexpr (G >= 16) v += dpp_f64<DPP_ROW_MIRROR>(0.0, v);
    if constexpr (G >= 32) {
        const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), 0x401F);  // xor 16 (bitmask mode)
        const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), 0x401F);
        v += __hiloint2double(hi, lo);
    }
    if constexpr (G >= 64) {
        const double lo_half = readlane_f64(v, 0), hi_half = readlane_f64(v, 32);
        v = lo_half + hi_half;
    }
    return v;
}

}  // namespace dev
}  // namespace axspmv
