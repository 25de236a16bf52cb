// Device helpers and the timing loop.
#include <algorithm>
#include <functional>
#include <vector>

#include "plan.hpp"

using namespace axspmv;

extern "C" int axspmv_dev_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(AXSPMV_E_INVALID, "dev_alloc: NULL");
    *ptr = nullptr;
    const hipError_t e = hipMalloc(ptr, bytes ? bytes : 256);
    return e == hipSuccess ? AXSPMV_OK : fail(AXSPMV_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
}

extern "C" int axspmv_dev_free(void *ptr) {
    const hipError_t e = hipFree(ptr);
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "hipFree");
}

extern "C" int axspmv_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return AXSPMV_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "memcpy h2d");
}

extern "C" int axspmv_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return AXSPMV_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "memcpy d2h");
}

extern "C" int axspmv_stream_sync(void *stream) {
    const hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "stream sync");
}

extern "C" int axspmv_device_sync(void) {
    const hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "device sync");
}

extern "C" int axspmv_device_count(int *count) {
    if (!count) return fail(AXSPMV_E_INVALID, "device_count: NULL");
    *count = 0;
    if (hipGetDeviceCount(count) != hipSuccess) *count = 0;
    return AXSPMV_OK;
}

extern "C" int axspmv_copy_f64(const double *d_src, double *d_dst, size_t n, void *stream) {
    const hipError_t e = launch::copy_f64(d_src, d_dst, n, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "copy launch");
}

extern "C" int axspmv_x_refresh(const uint32_t *d_map, uint64_t nslots, uint32_t hw, uint32_t nx, const double *d_x,
                                double *d_ax, void *stream) {
    if (!is_pow2(hw) || hw > 1024) return fail(AXSPMV_E_INVALID, "x_refresh: hw must be a power of two <= 1024, got %u", hw);
    if (nslots % hw) return fail(AXSPMV_E_INVALID, "x_refresh: %llu slots are not whole groups of %u",
                                 (unsigned long long)nslots, hw);
    if (nslots && (!d_map || !d_ax || (nx && !d_x))) return fail(AXSPMV_E_INVALID, "x_refresh: NULL argument");
    const hipError_t e = launch::x_refresh(d_map, nslots, ilog2(hw), nx, d_x, d_ax, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "x_refresh launch");
}

extern "C" int axspmv_wave_selftest(uint32_t op, const double *d_v, const uint32_t *d_h, double *d_ov,
                                    uint32_t *d_oh, uint64_t n, void *stream) {
    if (op > AXSPMV_WAVE_GROUP_SUM_64) return fail(AXSPMV_E_INVALID, "wave_selftest: unknown op %u", op);
    if (n % 64) return fail(AXSPMV_E_INVALID, "wave_selftest: %llu elements are not whole waves of 64",
                            (unsigned long long)n);
    if (n && (!d_v || !d_h || !d_ov || !d_oh)) return fail(AXSPMV_E_INVALID, "wave_selftest: NULL argument");
    const hipError_t e = launch::wave_selftest(op, d_v, d_h, d_ov, d_oh, n, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "wave_selftest launch");
}

namespace {

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

void summarize(std::vector<double> &ms, double scale, axspmv_timing *out) {
    for (auto &v : ms) v *= scale;
    double sum = 0.0;
    for (double v : ms) sum += v;
    std::vector<double> s = ms;
    std::sort(s.begin(), s.end());
    out->mean_ms = sum / double(ms.size());
    out->median_ms = s.size() % 2 ? s[s.size() / 2] : 0.5 * (s[s.size() / 2 - 1] + s[s.size() / 2]);
    out->min_ms = s.front();
    out->max_ms = s.back();
}

}  // namespace

// Timing loop.  EVENTS: mean of per-launch event times as in the reference
// (cudaSpmv.cu:563-581), plus the median.  GRAPH: hipGraph replays of
// `graph_batch` operations each, reported per operation.  REGION: one event
// pair around `iters` back-to-back launches.
int axspmv::time_launches(const std::function<hipError_t(hipStream_t)> &fn, uint32_t iters, uint32_t warmup,
                          uint32_t mode, uint32_t graph_batch, void *stream, axspmv_timing *out) {
    if (!out || iters == 0 || (mode > AXSPMV_TIME_REGION && mode != AXSPMV_TIME_COLD))
        return fail(AXSPMV_E_INVALID, "timing: bad argument");
    struct MainFromMean {
        axspmv_timing *o;
        ~MainFromMean() { o->main_ms = o->mean_ms; }
    } fill_main{out};
    std::memset(out, 0, sizeof *out);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipStream_t own = nullptr;
    Events ev;
    if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess)
        return fail(AXSPMV_E_HIP, "timing: event creation failed");
    std::vector<double> ms;
    ms.reserve(iters);
    out->mode = mode;
    out->iters = iters;
    if (mode != AXSPMV_TIME_GRAPH) {
        for (uint32_t i = 0; i < warmup; ++i)
            if (fn(s) != hipSuccess) return fail(AXSPMV_E_HIP, "timing: launch failed");
        if (mode == AXSPMV_TIME_COLD) {
            // Reading 512 MB (twice the Infinity Cache, far beyond the L2s)
            // evicts the operation's data without leaving dirty lines whose
            // write-back would land inside the timed launch.
            const size_t nflush = size_t(64) << 20;
            void *flush = nullptr;
            if (hipMalloc(&flush, nflush * sizeof(double) + 256) != hipSuccess)
                return fail(AXSPMV_E_NOMEM, "timing: cache-flush buffer allocation failed");
            hipError_t e = launch::fill_f64(static_cast<double *>(flush), nflush, 0.0, s);
            double *sink = static_cast<double *>(flush) + nflush;
            for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) {
                e = launch::touch(static_cast<const double *>(flush), nflush, sink, s);
                if (e == hipSuccess) e = hipEventRecord(ev.a, s);
                if (e == hipSuccess) e = fn(s);
                if (e == hipSuccess) e = hipEventRecord(ev.b, s);
                if (e == hipSuccess) e = hipEventSynchronize(ev.b);
                float t = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&t, ev.a, ev.b);
                ms.push_back(t);
            }
            (void)hipFree(flush);
            if (e != hipSuccess) return hip_fail(e, "timing (cold)");
            summarize(ms, 1.0, out);
            return AXSPMV_OK;
        }
        if (mode == AXSPMV_TIME_REGION) {
            hipError_t e = hipStreamSynchronize(s);
            if (e == hipSuccess) e = hipEventRecord(ev.a, s);
            for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) e = fn(s);
            if (e == hipSuccess) e = hipEventRecord(ev.b, s);
            if (e == hipSuccess) e = hipEventSynchronize(ev.b);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, ev.a, ev.b);
            if (e != hipSuccess) return hip_fail(e, "timing (region)");
            out->mean_ms = out->median_ms = out->min_ms = out->max_ms = double(t) / iters;
            return AXSPMV_OK;
        }
        for (uint32_t i = 0; i < iters; ++i) {
            hipError_t e = hipEventRecord(ev.a, s);
            if (e == hipSuccess) e = fn(s);
            if (e == hipSuccess) e = hipEventRecord(ev.b, s);
            if (e == hipSuccess) e = hipEventSynchronize(ev.b);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, ev.a, ev.b);
            if (e != hipSuccess) return hip_fail(e, "timing");
            ms.push_back(t);
        }
        summarize(ms, 1.0, out);
        return AXSPMV_OK;
    }
    if (graph_batch == 0) graph_batch = 1;
    if (!s) {
        if (hipStreamCreate(&own) != hipSuccess) return fail(AXSPMV_E_HIP, "timing: stream creation failed");
        s = own;
    }
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    for (uint32_t i = 0; i < graph_batch && e == hipSuccess; ++i) e = fn(s);
    const hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    for (uint32_t i = 0; i < warmup && e == hipSuccess; ++i) e = hipGraphLaunch(ge, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) {
        e = hipEventRecord(ev.a, s);
        if (e == hipSuccess) e = hipGraphLaunch(ge, s);
        if (e == hipSuccess) e = hipEventRecord(ev.b, s);
        if (e == hipSuccess) e = hipEventSynchronize(ev.b);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ev.a, ev.b);
        ms.push_back(t);
    }
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (own) (void)hipStreamDestroy(own);
    if (e != hipSuccess) return hip_fail(e, "timing (graph)");
    summarize(ms, 1.0 / graph_batch, out);
    return AXSPMV_OK;
}

namespace {

// AXSPMV_TIME_KERNEL: back-to-back SpMVs; event k[i] before SpMV i, m[i]
// between its two launches (if the path has two: plans with column slices),
// one more after the last.  main = m[i] .. k[i+1] after a clear of y, k[i] ..
// m[i] before a merge of slice partials.
int time_kernel(axspmv_plan *p, const axspmv_operands &o, uint32_t iters, uint32_t warmup, hipStream_t s,
                axspmv_timing *out) {
    if (iters > 4096) return fail(AXSPMV_E_INVALID, "timing: AXSPMV_TIME_KERNEL takes at most 4096 iterations");
    std::memset(out, 0, sizeof *out);
    out->mode = AXSPMV_TIME_KERNEL;
    out->iters = iters;
    std::vector<hipEvent_t> k(iters + 1, nullptr), m(iters, nullptr);
    struct Guard {
        std::vector<hipEvent_t> &a, &b;
        axspmv_plan *p;
        ~Guard() {
            p->mid = nullptr;
            for (hipEvent_t e : a)
                if (e) (void)hipEventDestroy(e);
            for (hipEvent_t e : b)
                if (e) (void)hipEventDestroy(e);
        }
    } guard{k, m, p};
    // Markers inside the region skip the system-scope fence (an L2 write-back
    // and invalidate that would slow the next SpMV); the closing event keeps
    // it, so its completion covers all the work.
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) {
        e = hipEventCreateWithFlags(&k[i], hipEventDisableSystemFence);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&m[i], hipEventDisableSystemFence);
    }
    if (e == hipSuccess) e = hipEventCreate(&k[iters]);
    if (e != hipSuccess) return hip_fail(e, "timing: event creation");
    // Sliced plans record a marker between their two launches: after the
    // clear of y (main kernel = what follows it) or, with slice partials,
    // before the merge (main kernel = what precedes it).
    const bool marks = (p->kind == Plan::RCT && p->rct.S > 1) || p->kind == Plan::AXT_UNC_H1_SLICED;
    const bool before_merge = (p->kind == Plan::RCT && p->rct.part) || (p->kind == Plan::AXT_UNC_H1_SLICED && p->d_part);
    for (uint32_t i = 0; i < warmup && e == hipSuccess; ++i) e = plan_launch(p, o.x, o.y, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) {
        e = hipEventRecord(k[i], s);
        p->mid = marks ? m[i] : nullptr;
        if (e == hipSuccess) e = plan_launch(p, o.x, o.y, s);
    }
    p->mid = nullptr;
    if (e == hipSuccess) e = hipEventRecord(k[iters], s);
    if (e == hipSuccess) e = hipEventSynchronize(k[iters]);
    float tot = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&tot, k[0], k[iters]);
    double main_sum = 0.0;
    std::vector<double> per(iters);  // SpMV i: k[i] .. k[i+1]
    for (uint32_t i = 0; i < iters && e == hipSuccess; ++i) {
        float t = 0.f;
        e = marks && before_merge ? hipEventElapsedTime(&t, k[i], m[i])
                                  : hipEventElapsedTime(&t, marks ? m[i] : k[i], k[i + 1]);
        main_sum += t;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, k[i], k[i + 1]);
        per[i] = t;
    }
    if (e != hipSuccess) return hip_fail(e, "timing (kernel)");
    summarize(per, 1.0, out);  // median, min and max of the SpMVs in the region
    out->mean_ms = double(tot) / iters;
    out->main_ms = main_sum / iters;
    return AXSPMV_OK;
}

}  // namespace

extern "C" int axspmv_time_op(axspmv_plan *p, uint32_t op, const axspmv_operands *ops, uint32_t iters,
                              uint32_t warmup, uint32_t mode, uint32_t graph_batch, void *stream, axspmv_timing *out) {
    if (!p || !ops) return fail(AXSPMV_E_INVALID, "time_op: NULL argument");
    AX_TRY(plan_op_check(p, op, *ops));
    AX_TRY(plan_op_prepare(p, op, *ops));
    const axspmv_operands o = *ops;
    if (mode == AXSPMV_TIME_KERNEL) {
        if (op != AXSPMV_OP_SPMV) return fail(AXSPMV_E_INVALID, "timing: AXSPMV_TIME_KERNEL times SpMV only");
        if (!out || iters == 0) return fail(AXSPMV_E_INVALID, "timing: bad argument");
        return time_kernel(p, o, iters, warmup, static_cast<hipStream_t>(stream), out);
    }
    return time_launches([p, op, o](hipStream_t st) { return plan_op(p, op, o, st); }, iters, warmup, mode,
                         graph_batch, stream, out);
}

extern "C" int axspmv_time_spmv(axspmv_plan *p, const double *d_x, double *d_y, uint32_t iters, uint32_t warmup,
                                uint32_t mode, uint32_t graph_batch, void *stream, axspmv_timing *out) {
    axspmv_operands o{};
    o.x = d_x;
    o.y = d_y;
    return axspmv_time_op(p, AXSPMV_OP_SPMV, &o, iters, warmup, mode, graph_batch, stream, out);
}

// ---------------------------------------------------------------- graphs
struct axspmv_graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    ~axspmv_graph() {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    }
};

namespace axspmv {

// Captures `repeats` calls of fn on a private stream into an executable graph.
int capture_graph(const std::function<hipError_t(hipStream_t)> &fn, uint32_t repeats, axspmv_graph **out) {
    if (!out || repeats == 0) return fail(AXSPMV_E_INVALID, "graph: bad argument");
    *out = nullptr;
    hipStream_t s = nullptr;
    if (hipStreamCreate(&s) != hipSuccess) return fail(AXSPMV_E_HIP, "graph: stream creation failed");
    auto gr = new axspmv_graph;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    for (uint32_t i = 0; i < repeats && e == hipSuccess; ++i) e = fn(s);
    const hipError_t e2 = hipStreamEndCapture(s, &gr->g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&gr->ge, gr->g, nullptr, nullptr, 0);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
        delete gr;
        return hip_fail(e, "graph capture");
    }
    *out = gr;
    return AXSPMV_OK;
}

}  // namespace axspmv

extern "C" int axspmv_plan_graph(axspmv_plan *p, uint32_t op, const axspmv_operands *ops, uint32_t repeats,
                                 axspmv_graph **out) {
    if (!p || !ops) return fail(AXSPMV_E_INVALID, "plan_graph: NULL argument");
    AX_TRY(plan_op_check(p, op, *ops));
    AX_TRY(plan_op_prepare(p, op, *ops));
    const axspmv_operands o = *ops;
    return capture_graph([p, op, o](hipStream_t st) { return plan_op(p, op, o, st); }, repeats, out);
}

extern "C" int axspmv_graph_launch(axspmv_graph *g, void *stream) {
    if (!g || !g->ge) return fail(AXSPMV_E_INVALID, "graph_launch: NULL graph");
    const hipError_t e = hipGraphLaunch(g->ge, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AXSPMV_OK : hip_fail(e, "graph launch");
}

extern "C" int axspmv_graph_destroy(axspmv_graph *g) {
    delete g;
    return AXSPMV_OK;
}
