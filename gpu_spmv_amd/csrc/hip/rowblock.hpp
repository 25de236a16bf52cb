// Row-aligned wave-block engine shared by the AXT UNC kernels and the
// batched persistent kernel.
//
// A wave owns a contiguous range of units [b.beg, b.end) that starts and ends
// on row boundaries, and walks it 64 units at a time.  Each lane sums its
// unit; lane sums are combined across lanes by a DPP segmented scan keyed on
// the row id; the last lane of each row segment stores y.  A segment still
// open at lane 63 is carried to the next 64 units in scalar registers.
#pragma once

#include "launch.hpp"
#include "wave64.hpp"

namespace axspmv {
namespace dev {
using axspmv::launch::Block;
using axspmv::launch::Window;

// x as the kernels read it: straight from HBM, or from a wave's LDS copy of
// the columns [cmin, cmin + width) its block touches.  Padding slots carry
// the column 0xffffffff and read 0 either way (outside every window).
struct XGlobal {
    const double *__restrict__ p;
    __device__ __forceinline__ double operator()(uint32_t c) const { return xpad(p, c); }
};
struct XWindow {
    const double *xw;
    uint32_t cmin, width;
    __device__ __forceinline__ double operator()(uint32_t c) const {
        const uint32_t i = c - cmin;
        return i < width ? xw[i] : 0.0;
    }
};

// LDS stores of this wave become visible to its own later loads (and loads
// finish before later stores), without a workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The wave copies x[cmin, cmin + width) into xw (width <= STAGE_COLS).
__device__ __forceinline__ XWindow stage_window(double *xw, const Window w, const double *__restrict__ x) {
    const uint32_t lane = threadIdx.x & 63;
    wave_lds_sync();  // the previous item's reads are done
    for (uint32_t i = lane; i < w.width; i += 64) xw[i] = x[w.cmin + i];
    wave_lds_sync();
    return XWindow{xw, w.cmin, w.width};
}

typedef double sf64x2 __attribute__((ext_vector_type(2)));
typedef float sf32x2 __attribute__((ext_vector_type(2)));
typedef float sf32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t su32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t su32x4 __attribute__((ext_vector_type(4)));

template <class VT, int K, bool NT = false>
struct VecLoad;
template <bool NT>
struct VecLoad<double, 4, NT> {
    __device__ static void get(const double *p, double *v) {
        const sf64x2 a = stream_load<NT>(reinterpret_cast<const sf64x2 *>(p)),
                     b = stream_load<NT>(reinterpret_cast<const sf64x2 *>(p) + 1);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
};
template <bool NT>
struct VecLoad<double, 2, NT> {
    __device__ static void get(const double *p, double *v) {
        const sf64x2 a = stream_load<NT>(reinterpret_cast<const sf64x2 *>(p));
        v[0] = a.x; v[1] = a.y;
    }
};
template <bool NT>
struct VecLoad<double, 1, NT> {
    __device__ static void get(const double *p, double *v) { v[0] = stream_load<NT>(p); }
};
template <bool NT>
struct VecLoad<float, 4, NT> {
    __device__ static void get(const float *p, float *v) {
        const sf32x4 a = stream_load<NT>(reinterpret_cast<const sf32x4 *>(p));
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
};
template <bool NT>
struct VecLoad<float, 2, NT> {
    __device__ static void get(const float *p, float *v) {
        const sf32x2 a = stream_load<NT>(reinterpret_cast<const sf32x2 *>(p));
        v[0] = a.x; v[1] = a.y;
    }
};
template <bool NT>
struct VecLoad<float, 1, NT> {
    __device__ static void get(const float *p, float *v) { v[0] = stream_load<NT>(p); }
};
template <int K, bool NT = false>
__device__ __forceinline__ void load_cols(const uint32_t *p, uint32_t *c) {
    if constexpr (K == 4) {
        const su32x4 a = stream_load<NT>(reinterpret_cast<const su32x4 *>(p));
        c[0] = a.x; c[1] = a.y; c[2] = a.z; c[3] = a.w;
    } else if constexpr (K == 2) {
        const su32x2 a = stream_load<NT>(reinterpret_cast<const su32x2 *>(p));
        c[0] = a.x; c[1] = a.y;
    } else {
        c[0] = stream_load<NT>(p);
    }
}

// th == 1: unit u covers plane slots [u*K, u*K + K), tile (u*K)/thw.  NT:
// the planes are streamed with non-temporal loads.
template <class VT, int K, bool NT = false>
struct UncH1Src {
    const VT *__restrict__ vp;
    const uint32_t *__restrict__ cp;
    const uint32_t *__restrict__ sec;
    uint32_t ushift;  // log2(thw / K)
    template <class XA>
    __device__ __forceinline__ double sum(uint32_t u, const XA &x) const {
        const size_t q = size_t(u) * K;
        VT v[K];
        uint32_t c[K];
        VecLoad<VT, K, NT>::get(vp + q, v);
        load_cols<K, NT>(cp + q, c);
        double xv[K];
#pragma unroll
        for (int j = 0; j < K; ++j) xv[j] = x(c[j]);
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) acc = fma(double(v[j]), xv[j], acc);
        return acc;
    }
    __device__ __forceinline__ uint32_t row(uint32_t u) const { return sec[u >> ushift]; }
    // As sum(), keeping the gathered x values.
    __device__ __forceinline__ double sum_keep(uint32_t u, const double *__restrict__ x, double *xv) const {
        const size_t q = size_t(u) * K;
        VT v[K];
        uint32_t c[K];
        VecLoad<VT, K, NT>::get(vp + q, v);
        load_cols<K, NT>(cp + q, c);
#pragma unroll
        for (int j = 0; j < K; ++j) xv[j] = xpad(x, c[j]);
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) acc = fma(double(v[j]), xv[j], acc);
        return acc;
    }
    __device__ __forceinline__ void gather_x(uint32_t u, const double *__restrict__ x, double *xv) const {
        uint32_t c[K];
        load_cols<K, NT>(cp + size_t(u) * K, c);
#pragma unroll
        for (int j = 0; j < K; ++j) xv[j] = xpad(x, c[j]);
    }
};

// th == 1 on the interleaved ax array: unit u covers value slots [u*K,
// u*K + K) of tile t = (u*K) / thw, stored at ax[t*2thw + (u*K) % thw] with
// their x values thw further on.
template <int K>
struct AxH1Src {
    const double *__restrict__ ax;
    const uint32_t *__restrict__ sec;
    uint32_t ushift;   // log2(thw / K)
    uint32_t thw_log;
    template <class XA>
    __device__ __forceinline__ double sum(uint32_t u, const XA &) const {
        const size_t q = size_t(u) * K;
        const size_t p = ((q >> thw_log) << (thw_log + 1)) + (q & ((size_t(1) << thw_log) - 1));
        double v[K], xv[K];
        VecLoad<double, K>::get(ax + p, v);
        VecLoad<double, K>::get(ax + p + (size_t(1) << thw_log), xv);
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) acc = fma(v[j], xv[j], acc);
        return acc;
    }
    __device__ __forceinline__ uint32_t row(uint32_t u) const { return sec[u >> ushift]; }
};

// th > 1: unit g is global tile column g; level f sits at
// (g/thw)*thw*th + f*thw + g%thw.  TH > 0 fixes th at compile time.  NT: the
// planes are streamed with non-temporal loads.
template <class VT, int TH, bool NT = false>
struct UncHSrc {
    const VT *__restrict__ vp;
    const uint32_t *__restrict__ cp;
    const uint32_t *__restrict__ sec;
    uint32_t thw_log, th;
    template <class XA>
    __device__ __forceinline__ double sum(uint32_t g, const XA &x) const {
        const uint32_t thw = 1u << thw_log;
        const uint32_t h = TH > 0 ? uint32_t(TH) : th;
        const size_t base = size_t(g >> thw_log) * (size_t(h) << thw_log) + (g & (thw - 1));
        double acc = 0.0;
        if constexpr (TH > 0) {
            double xv[TH];
            VT v[TH];
#pragma unroll
            for (int f = 0; f < TH; ++f) {
                v[f] = stream_load<NT>(vp + base + size_t(f) * thw);
                xv[f] = x(stream_load<NT>(cp + base + size_t(f) * thw));
            }
#pragma unroll
            for (int f = 0; f < TH; ++f) acc = fma(double(v[f]), xv[f], acc);
        } else {
#pragma unroll 4
            for (uint32_t f = 0; f < h; ++f)
                acc = fma(double(stream_load<NT>(vp + base + size_t(f) * thw)),
                          x(stream_load<NT>(cp + base + size_t(f) * thw)), acc);
        }
        return acc;
    }
    __device__ __forceinline__ uint32_t row(uint32_t g) const { return sec[g]; }
};

// ADD: the block holds the row sums of one column slice; they are added to
// y (cleared beforehand) with FP64 atomics instead of stored.
template <bool ADD>
__device__ __forceinline__ void put_row(double *__restrict__ y, uint32_t r, double v) {
    if constexpr (ADD)
        unsafeAtomicAdd(y + r, v);
    else
        y[r] = v;
}

template <class Src, class XA, bool ADD = false>
__device__ __forceinline__ void rowblock_process(const Src &src, const Block b, const XA &x, double *__restrict__ y) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t crow = NOROW;  // row of the segment still open at the end of the previous 64 units
    double carry = 0.0;
    for (uint32_t u0 = b.beg; u0 < b.end; u0 += 64) {
        const uint32_t u = u0 + lane;
        const bool act = u < b.end;
        double s = 0.0;
        uint32_t r = NOROW;
        if (act) {
            s = src.sum(u, x);
            r = src.row(u);
        }
        // The carried segment ends here if these units start with another row.
        const uint32_t r0 = readfirst_u32(r);
        if (crow != NOROW && r0 != crow) {
            if (lane == 0) put_row<ADD>(y, crow, carry);
            crow = NOROW;
            carry = 0.0;
        }
        uint32_t head = r != shift_up_u32(r, crow);
        double v = seg_scan_f64(s, head);
        if (!head) v += carry;  // still inside the carried row
        const bool last = u0 + 64 >= b.end;
        bool end = act && shift_down_u32(r, NOROW) != r;
        if (!last && lane == 63) end = false;
        if (end) put_row<ADD>(y, r, v);
        if (!last) {
            crow = readlane_u32(r, 63);
            carry = readlane_f64(v, 63);
        }
    }
}

}  // namespace dev
}  // namespace axspmv
