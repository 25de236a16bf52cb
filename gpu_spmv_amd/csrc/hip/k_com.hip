// AXT COM kernels (compacted: slots in CSR order, per-slot segment headers
// (row << log) | (segment length - 1) at every row start and every block /
// column start).
//
// The reference adds every segment into a zeroed y with FP64 atomics
// (gaxtch1, cudaSpmv.cu:2041-2097, through a block-wide scan and the
// difference S[i+o] - S[i] + v_i; gaxtch, cudaSpmv.cu:2174-2209, lane walking
// its column).  Here:
//
//  * a wave owns a fixed range of slots (whole bs blocks for th == 1, whole
//    tile columns for th > 1) and reduces it by row with a DPP segmented scan
//    keyed on the row id decoded from the headers (no scan differences, so no
//    cancellation);
//  * every row that starts and ends inside the wave is stored once;
//  * the first run of a wave that continues a row from the previous wave and
//    the last run that continues into the next wave are written to per-wave
//    partials, and a second small kernel sums the partials of those spanning
//    rows in wave order: y is fully overwritten, no atomics, deterministic.
//
// Lane layout.  th == 1: a lane holds 4 consecutive slots (vector loads of
// values, columns, headers).  th > 1: a lane holds one tile column, th
// consecutive slots at stride thw in memory (the levels of 64 neighbouring
// columns are contiguous, so loads coalesce).
#include "launch.hpp"
#include "wave64.hpp"

namespace {
using namespace axspmv::dev;
using axspmv::launch::ComArgs;
using axspmv::launch::ComWave;

// Per-lane summary of a run of slots: first run (row fr, sum fs), last run
// (lr, ls); brk = more than one run in the lane.  Middle runs are complete
// rows and are stored by the lane itself.
struct LaneSum {
    uint32_t fr, lr;
    double fs, ls;
    bool brk, any;
};

// Handles one run that ends: stores y, or the wave's head partial when the
// run is the wave's first run and continues a row from the previous wave.
__device__ __forceinline__ void emit(uint32_t row, double v, uint32_t first_row, bool cont_in, double *y, double *head) {
    if (cont_in && row == first_row)
        *head = v;
    else
        y[row] = v;
}

// Combines the lane summaries of one chunk of the wave's range.  (crow, cval)
// is the run still open at the end of the previous chunk; it is updated.
__device__ __forceinline__ void combine(const LaneSum &L, bool last_chunk, uint32_t &crow, double &cval,
                                        uint32_t first_row, bool cont_in, bool cont_out, double *y, double *head,
                                        double *tail) {
    const uint32_t lane = lane_id();
    const uint32_t fr = L.any ? L.fr : NOROW, lr = L.any ? L.lr : NOROW;
    // The carried run ends where this chunk starts another row.
    const uint32_t fr0 = readfirst_u32(fr);
    if (crow != NOROW && fr0 != crow) {
        if (lane == 0) emit(crow, cval, first_row, cont_in, y, head);
        crow = NOROW;
        cval = 0.0;
    }
    const uint32_t prev_lr = shift_up_u32(lr, crow);
    uint32_t t = (L.brk || fr != prev_lr) ? 1u : 0u;
    double c = seg_scan_f64(L.any ? L.ls : 0.0, t);
    if (!t) c += cval;  // trailing run reaches back into the carried run
    const double c_prev = shift_up_f64(c, cval);
    if (L.any && L.brk) {
        // first run of the lane ends inside the lane
        const double tot = L.fs + (fr == prev_lr ? c_prev : 0.0);
        emit(fr, tot, first_row, cont_in, y, head);
    }
    // trailing run: ends here if the next lane starts another row
    const uint32_t next_fr = shift_down_u32(fr, NOROW);
    const uint64_t act = __ballot(L.any);  // active lanes are contiguous from lane 0
    const uint32_t hi = act ? 63 - __builtin_clzll(act) : 0;
    if (L.any) {
        if (lane == hi) {
            if (last_chunk) {
                if (cont_out)
                    *tail = c;
                else
                    emit(lr, c, first_row, cont_in, y, head);
                if (cont_out && cont_in && lr == first_row) *head = c;  // a wave inside one row
            }
        } else if (next_fr != lr) {
            emit(lr, c, first_row, cont_in, y, head);
        }
    }
    if (!last_chunk) {
        crow = readlane_u32(lr, hi);
        cval = readlane_f64(c, hi);
    }
}

// th == 1: lane holds slots [p0, p0 + 4) of the CSR-ordered planes.  NT:
// the planes are streamed with non-temporal loads.
template <class VT, bool NT>
__device__ __forceinline__ LaneSum lane_com_h1(const ComArgs &a, uint32_t p0, uint32_t pend, uint32_t in_row,
                                               const double *__restrict__ x, double *__restrict__ y,
                                               uint32_t first_row, bool cont_in, double *head) {
    const VT *vp = static_cast<const VT *>(a.vplane);
    LaneSum L{NOROW, NOROW, 0.0, 0.0, false, false};
    double prod[4];
    uint32_t rows[4];
    bool act[4];
    uint32_t hdr[4], col[4];
    VT v[4];
    if (p0 + 4 <= pend) {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 h4 = stream_load<NT>(reinterpret_cast<const u32x4 *>(a.hdr + p0));
        const u32x4 c4 = stream_load<NT>(reinterpret_cast<const u32x4 *>(a.cplane + p0));
        hdr[0] = h4.x; hdr[1] = h4.y; hdr[2] = h4.z; hdr[3] = h4.w;
        col[0] = c4.x; col[1] = c4.y; col[2] = c4.z; col[3] = c4.w;
#pragma unroll
        for (int j = 0; j < 4; ++j) { v[j] = stream_load<NT>(vp + p0 + j); act[j] = true; }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            act[j] = p0 + j < pend;
            hdr[j] = act[j] ? stream_load<NT>(a.hdr + p0 + j) : 0u;
            col[j] = act[j] ? stream_load<NT>(a.cplane + p0 + j) : 0u;
            v[j] = act[j] ? stream_load<NT>(vp + p0 + j) : VT(0);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) prod[j] = act[j] ? double(v[j]) * x[col[j]] : 0.0;
    uint32_t cur = in_row;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const bool head_slot = act[j] && (((p0 + j) & (a.bs - 1)) == 0 || hdr[j] != 0);
        if (head_slot) cur = hdr[j] >> a.log;
        rows[j] = cur;
    }
    // runs inside the lane
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (!act[j]) break;
        if (!L.any) {
            L.any = true;
            L.fr = L.lr = rows[j];
            L.fs = L.ls = prod[j];
        } else if (rows[j] == L.lr) {
            L.ls += prod[j];
            if (!L.brk) L.fs = L.ls;
        } else {
            if (L.brk) emit(L.lr, L.ls, first_row, cont_in, y, head);  // a middle run: complete
            L.brk = true;
            L.lr = rows[j];
            L.ls = prod[j];
        }
    }
    return L;
}

template <class VT, bool NT>
__global__ void __launch_bounds__(256) com_h1_kernel(ComArgs a, const double *__restrict__ x, double *__restrict__ y) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= a.nwaves) return;
    const ComWave w = a.waves[wave];
    const bool cont_in = w.flags & 1, cont_out = w.flags & 2;
    double *head = a.head + wave, *tail = a.tail + wave;
    uint32_t crow = NOROW, crow_hdr = 0;  // crow_hdr: row of the last decoded header + 1
    double cval = 0.0;
    const uint32_t first_row = a.hdr[w.beg] >> a.log;  // a wave range starts on a bs block
    for (uint32_t c0 = w.beg; c0 < w.end; c0 += 256) {
        const uint32_t p0 = c0 + lane * 4;
        // row inherited by slots before the lane's first header: max over
        // earlier lanes' headers (rows grow along the CSR order)
        uint32_t mymax = 0;
        if (p0 < w.end) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t p = p0 + j;
                if (p < w.end) {
                    const uint32_t h = a.hdr[p];
                    if ((p & (a.bs - 1)) == 0 || h != 0) mymax = max(mymax, (h >> a.log) + 1);
                }
            }
        }
        const uint32_t incl = scan_max_u32(mymax);
        uint32_t excl = shift_up_u32(incl, 0u);
        excl = max(excl, crow_hdr);
        const uint32_t in_row = excl ? excl - 1 : 0;
        LaneSum L{NOROW, NOROW, 0.0, 0.0, false, false};
        if (p0 < w.end) L = lane_com_h1<VT, NT>(a, p0, w.end, in_row, x, y, first_row, cont_in, head);
        const bool last = c0 + 256 >= w.end;
        combine(L, last, crow, cval, first_row, cont_in, cont_out, y, head, tail);
        crow_hdr = max(crow_hdr, readlane_u32(incl, 63));
    }
}

// th > 1: lane holds global tile column g (slots g*th .. g*th + th - 1).
template <class VT, int TH, bool NT>
__global__ void __launch_bounds__(256) com_h_kernel(ComArgs a, const double *__restrict__ x, double *__restrict__ y) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= a.nwaves) return;
    const VT *vp = static_cast<const VT *>(a.vplane);
    const ComWave w = a.waves[wave];  // beg/end in columns
    const bool cont_in = w.flags & 1, cont_out = w.flags & 2;
    double *head = a.head + wave, *tail = a.tail + wave;
    const uint32_t thw = a.thw, thw_log = __builtin_ctz(thw);
    const uint32_t th = TH > 0 ? uint32_t(TH) : a.th;
    uint32_t crow = NOROW;
    double cval = 0.0;
    const uint32_t first_row = a.hdr[size_t(w.beg >> thw_log) * th * thw + (w.beg & (thw - 1))] >> a.log;
    for (uint32_t g0 = w.beg; g0 < w.end; g0 += 64) {
        const uint32_t g = g0 + lane;
        LaneSum L{NOROW, NOROW, 0.0, 0.0, false, false};
        if (g < w.end) {
            const size_t base = size_t(g >> thw_log) * th * thw + (g & (thw - 1));
            const uint64_t p0 = uint64_t(g) * th;
            // levels holding real entries (the last tile column may be short)
            const uint32_t nf = p0 >= a.nnz ? 0u : uint32_t(min<uint64_t>(th, a.nnz - p0));
            constexpr uint32_t NB = TH > 0 ? uint32_t(TH) : 1u;
            double pr[NB];
            uint32_t hd[NB];
            if constexpr (TH > 0) {
                // issue every level's loads before the first use
                uint32_t cc[NB];
                VT vv[NB];
#pragma unroll
                for (uint32_t f = 0; f < NB; ++f) {
                    const size_t q = base + size_t(f) * thw;
                    const bool ok = f < nf;
                    hd[f] = ok ? stream_load<NT>(a.hdr + q) : 0u;
                    cc[f] = ok ? stream_load<NT>(a.cplane + q) : 0u;
                    vv[f] = ok ? stream_load<NT>(vp + q) : VT(0);
                }
#pragma unroll
                for (uint32_t f = 0; f < NB; ++f) pr[f] = f < nf ? double(vv[f]) * x[cc[f]] : 0.0;
            }
            uint32_t cur = 0;
            for (uint32_t f = 0; f < nf; ++f) {
                uint32_t h;
                double prod;
                if constexpr (TH > 0) {
                    h = hd[f];
                    prod = pr[f];
                } else {
                    const size_t q = base + size_t(f) * thw;
                    h = stream_load<NT>(a.hdr + q);
                    prod = double(stream_load<NT>(vp + q)) * x[stream_load<NT>(a.cplane + q)];
                }
                if (f == 0 || h != 0) cur = h >> a.log;
                if (!L.any) {
                    L.any = true;
                    L.fr = L.lr = cur;
                    L.fs = L.ls = prod;
                } else if (cur == L.lr) {
                    L.ls += prod;
                    if (!L.brk) L.fs = L.ls;
                } else {
                    if (L.brk) emit(L.lr, L.ls, first_row, cont_in, y, head);
                    L.brk = true;
                    L.lr = cur;
                    L.ls = prod;
                }
            }
        }
        const bool last = g0 + 64 >= w.end;
        combine(L, last, crow, cval, first_row, cont_in, cont_out, y, head, tail);
    }
}

// Rows that span waves: y[row] = tail[w0] + head[w0+1] + ... + head[w1].
__global__ void __launch_bounds__(256) com_fixup(const uint32_t *__restrict__ span, uint32_t nspan,
                                                 const double *__restrict__ head, const double *__restrict__ tail,
                                                 double *__restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nspan) return;
    const uint32_t row = span[3 * i], w0 = span[3 * i + 1], w1 = span[3 * i + 2];
    double acc = tail[w0];
    for (uint32_t w = w0 + 1; w <= w1; ++w) acc += head[w];
    y[row] = acc;
}

template <class VT, bool NT>
void com_launch(const ComArgs &a, const double *x, double *y, hipStream_t s) {
    const unsigned grid = (a.nwaves + 3) / 4;
    if (a.th == 1) {
        com_h1_kernel<VT, NT><<<grid, 256, 0, s>>>(a, x, y);
        return;
    }
    switch (a.th) {
#define AX_TH(T) \
    case T: com_h_kernel<VT, T, NT><<<grid, 256, 0, s>>>(a, x, y); break;
        AX_TH(2)
        AX_TH(4)
        AX_TH(8)
        AX_TH(16)
#undef AX_TH
        default: com_h_kernel<VT, 0, NT><<<grid, 256, 0, s>>>(a, x, y);
    }
}

template <class VT>
hipError_t com_vt(const ComArgs &a, const double *x, double *y, hipStream_t s) {
    if (a.nwaves) {
        if (a.nt)
            com_launch<VT, true>(a, x, y, s);
        else
            com_launch<VT, false>(a, x, y, s);
    }
    if (a.nempty) {
        // rows without entries never appear in a header
        const hipError_t e = axspmv::launch::fill_rows(a.empty, a.nempty, y, s);
        if (e != hipSuccess) return e;
    }
    if (a.nspan) com_fixup<<<(a.nspan + 255) / 256, 256, 0, s>>>(a.span, a.nspan, a.head, a.tail, y);
    return hipGetLastError();
}

}  // namespace

namespace axspmv {
namespace launch {

hipError_t axt_com(const ComArgs &a, const double *x, double *y, hipStream_t s) {
    return a.vtype == 1 ? com_vt<float>(a, x, y, s) : com_vt<double>(a, x, y, s);
}

}  // namespace launch
}  // namespace axspmv
