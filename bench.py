"""SpMV benchmark on one MI355X.

Workload of the headline line (one step = one SpMV y = A x, inputs resident
in HBM): the synthetic power-law matrix, n = 2^20 rows, row lengths Zipf(alpha
= 2) on {1..2048} rescaled to mean 20 (nnz = 20,970,465), uniform sorted
distinct columns, values N(0,1), x ~ U[-1,1], seed 1 -- FP64.

Prints one JSON line: GFLOP/s of the flagship plan (2*nnz per SpMV, host
clock over the timed region), its roofline against 8 TB/s of HBM
(algorithmic bytes per SpMV / the whole SpMV's time from HIP events, with
the tile kernel alone beside it, the spread over the region and the
cold-cache time), the CPU baseline (the reference's own CSR product, ncsr,
built from its sources into oracle/_ref, else the oracle's restatement of
it, on this job's host cores), and, per configuration, the reference's two
result tables: formats (name, memory footprint, occupancy, conversion time
on the host and on the device) and kernels (time, GFLOP/s, GB/s, fraction
of the roofline, ordering time, errAbs / errRel / rowInd as the reference's
get_errors computes them and the signed error).  By default every
configuration runs: the power-law matrix (every format, forward + adjoint,
SpMM at k = 8 / 32 / 64 in FP64 and FP32 values with the error against the
FP64 oracle, conversion times, rocSPARSE), the 256^2 Laplacian
(graph-timed), the permuted 27-point stencil on 128^3, forward + adjoint on
its structure, and the 256-matrix banded batch in one launch.

  python bench.py [--steps K] [--warmup W] [--formats flagship|all]
                  [--suite all|headline]
  python bench.py --calibrate      copy bandwidth and the x-footprint sweep
  python bench.py --pmc-probe [stencil|spmm|stencil_rcm|batch256|fwd_adj_rcm]
                                   a 256 MB copy then 5 SpMVs of each flagship
                                   kernel on the power-law matrix (or on the
                                   permuted 27-point stencil), or 3 SpMMs at
                                   k = 8, 32, 64, for rocprofv3 --pmc runs
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpu_spmv_amd as ax  # noqa: E402
from gpu_spmv_amd import _lib as L  # noqa: E402
from gpu_spmv_amd import workloads as W  # noqa: E402
from gpu_spmv_amd.core import TIME_COLD, TIME_GRAPH, TIME_KERNEL, TIME_REGION  # noqa: E402

PEAK_HBM_GBS = 8000.0
FLAGSHIP = ("rct_auto", ax.FMT_RCT, dict())
FLAGSHIP_AXT = ("axt_unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1))
# the reference's AXT UNC th = 1 layout, one per column slice of 2 MB of x
FLAGSHIP_AXT_SLICED = ("axt_unc_h1_thw2_s4", ax.FMT_AXT, dict(thw=2, slices=4))
OTHERS = [
    FLAGSHIP_AXT,
    FLAGSHIP_AXT_SLICED,
    ("axt_unc_h1_thw2_s4_det", ax.FMT_AXT, dict(thw=2, slices=4, flags=ax.PLAN_DETERMINISTIC)),
    ("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
    ("csr_vector", ax.FMT_CSR, dict(kernel=ax.KER_STREAM)),
    ("axt_unc_h1_thw8", ax.FMT_AXT, dict(thw=8, th=1)),
    ("axt_unc_h1_thw16", ax.FMT_AXT, dict(thw=16, th=1)),
    ("axt_unc_h4_thw64", ax.FMT_AXT, dict(thw=64, th=4)),
    ("axt_unc_h8_thw64", ax.FMT_AXT, dict(thw=64, th=8)),
    ("axt_com_h1_thw64_bs256", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
    ("axt_com_h4_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4)),
    ("axt_com_h8_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=8)),
    ("k1_c64", ax.FMT_K1, dict(chunk=64)),
    ("k1_c64_natural", ax.FMT_K1, dict(chunk=64, reorder=0)),
    ("axc_hbs4", ax.FMT_AXC, dict(hbs=4)),
    ("axc_hbs64", ax.FMT_AXC, dict(hbs=64)),
    ("rct_r8192_s4", ax.FMT_RCT, dict(tile_rows=8192, slices=4)),
    ("rct_r8192_s8", ax.FMT_RCT, dict(tile_rows=8192, slices=8)),
    ("rct_r16384_s2", ax.FMT_RCT, dict(tile_rows=16384, slices=2)),
    ("axt_unc_h1_thw4_ax", ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_AX_LAYOUT)),
    ("rct_auto_f32val", ax.FMT_RCT, dict(vtype=ax.F32)),
    ("axt_unc_h1_thw4_f32val", ax.FMT_AXT, dict(thw=4, vtype=ax.F32)),
]
# C3's rows with AXSPMV_PLAN_REORDER (the plan of B = Q A Q^T, Q the reverse
# Cuthill-McKee order; x and y permuted inside the timed SpMV)
RCM = ax.PLAN_REORDER
REORDERED = [
    ("rct_auto_rcm", ax.FMT_RCT, dict(flags=RCM)),
    ("rct_r8192_s1_rcm", ax.FMT_RCT, dict(tile_rows=8192, slices=1, flags=RCM)),
    ("axt_unc_h1_thw4_rcm", ax.FMT_AXT, dict(thw=4, flags=RCM)),
    ("axt_unc_h4_thw64_rcm", ax.FMT_AXT, dict(thw=64, th=4, flags=RCM)),
    ("axt_com_h4_thw64_rcm", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4, flags=RCM)),
    ("k1_c64_rcm", ax.FMT_K1, dict(chunk=64, flags=RCM)),
    ("axc_hbs4_rcm", ax.FMT_AXC, dict(hbs=4, flags=RCM)),
    ("csr_vector_rcm", ax.FMT_CSR, dict(kernel=ax.KER_STREAM, flags=RCM)),
]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err(ref, y):
    den = float(np.max(np.abs(ref))) if ref.size else 0.0
    return float(np.max(np.abs(y - ref)) / den) if den else float(np.max(np.abs(y - ref)))


def omp_threads():
    return int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1)


REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_dump_c64")


CMP_LIB = os.path.join(ROOT, "gpu_spmv_amd", "librocsparse_cmp.so")
ROCSPARSE_ALGS = ("adaptive", "rowsplit", "lrb", "nnzsplit")


def rocsparse_rows(m, dx, yref, iters: int = 50):
    """rocSPARSE's CSR SpMV on the same matrix and x, each of its CSR
    algorithms (the counterpart of the reference's cuSPARSE row,
    cudaSpmv.cu:629-704): a comparator built into librocsparse_cmp.so,
    never a library path.  None when that library is not built.  Call it
    from a process that has not imported torch (torch ships its own
    rocSPARSE build, which librocsparse.so.1 could then resolve to)."""
    import ctypes as C
    if not os.path.exists(CMP_LIB):
        return None
    lib = C.CDLL(CMP_LIB)
    u32, vp = C.c_uint32, C.c_void_p
    lib.cmp_rocsparse_spmv.argtypes = [u32, u32, u32, vp, vp, vp, vp, vp, C.c_int, u32,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.cmp_rocsparse_spmv.restype = C.c_int
    lib.cmp_last_error.restype = C.c_char_p
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    drow, dcol, dval = (ax.DeviceBuffer.from_array(a) for a in (m.row, m.col, m.val))
    dy = ax.DeviceBuffer(8 * n)
    alg = 12.0 * nnz + 4.0 * (n + 1) + 8.0 * nc + 8.0 * n
    rows = []
    for i, name in enumerate(ROCSPARSE_ALGS):
        an, t, api = C.c_double(), C.c_double(), C.c_int()
        rc = lib.cmp_rocsparse_spmv(n, nc, nnz, drow.ptr, dcol.ptr, dval.ptr, dx.ptr, dy.ptr, i, iters,
                                    C.byref(an), C.byref(t), C.byref(api))
        if rc:
            rows.append(dict(label=f"rocsparse_{name}", error=lib.cmp_last_error().decode()))
            continue
        y = dy.download(np.float64, n)
        sec = t.value * 1e-3
        rows.append(dict(label=f"rocsparse_{name}", us=round(sec * 1e6, 2), gflops=round(2.0 * nnz / sec / 1e9, 2),
                         alg_gbs=round(alg / sec / 1e9, 1), frac=round(alg / sec / 1e9 / PEAK_HBM_GBS, 4),
                         analysis_ms=round(an.value, 2), api=f"rocsparse v{api.value}",
                         rel_err=rel_err(yref, y) if yref is not None else None))
        log(f"rocsparse: {rows[-1]}")
    return rows


def cpu_info() -> dict:
    """The host's CPU model, its physical cores and logical CPUs, and the
    CPUs this process may run on (its share of the host)."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
    except OSError:
        pass
    if phys is not None and core is not None:
        cores.add((phys, core))
    try:
        share = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        share = os.cpu_count() or 1
    return dict(cpu=model, host_physical_cores=len(cores) or None, host_logical_cpus=os.cpu_count(),
                cpus_available_to_this_job=share)


def cpu_baseline_reference(m, x, seconds: float):
    """The reference's own ncsr (its host code compiled in place into
    oracle/_ref), OpenMP over the host threads, repeated for `seconds`; None
    when that binary is not built or the matrix is not square (it reads x of
    nrows entries)."""
    import subprocess
    import tempfile
    if not os.path.exists(REF_BIN) or m.nrows != m.ncols:
        return None
    threads = omp_threads()
    with tempfile.TemporaryDirectory() as tmp:
        mat, xf, out = (os.path.join(tmp, f) for f in ("m.bin", "x.bin", "o"))
        m.write(mat)
        np.ascontiguousarray(x[:m.nrows], dtype=np.float64).tofile(xf)
        # the OpenMP settings of the reference's launcher (jobLauncher.sh:171-173)
        env = dict(os.environ, OMP_WAIT_POLICY="active", OMP_DYNAMIC="false", OMP_PROC_BIND="true")
        try:
            subprocess.run([REF_BIN, mat, xf, out, "time", str(seconds), str(threads)], check=True,
                           stdout=subprocess.DEVNULL, timeout=seconds * 6 + 120, env=env)
        except (subprocess.SubprocessError, OSError) as e:
            log(f"cpu baseline: reference ncsr did not run ({e}); timing the oracle port")
            return None
        meta = dict(line.split(None, 1) for line in open(out + ".meta").read().splitlines() if line.strip())
        y = np.fromfile(out + ".y", np.float64)
    calls, el = int(meta["calls"]), float(meta["seconds"])
    per = el / calls
    alg = 12.0 * m.nnz + 4.0 * (m.nrows + 1) + 8.0 * m.ncols + 8.0 * m.nrows
    return y, dict(value=round(2.0 * m.nnz / per / 1e9, 3), unit="GFLOP/s", cores=threads, kind="reference",
                   sample=f"{calls} calls of the reference's ncsr (cudaSpmv.cu:451-466, OpenMP over rows, "
                          f"OMP_PROC_BIND=true, OMP_WAIT_POLICY=active, built from its sources into oracle/_ref) on the "
                          f"full matrix, "
                          f"{el:.1f} s",
                   ms_per_spmv=round(per * 1e3, 3), alg_gbs=round(alg / per / 1e9, 1), **cpu_info())


def cpu_baseline(m, x, seconds: float):
    """The reference's ncsr where oracle/_ref holds it, else the oracle's
    restatement of it, on the host threads."""
    ref = cpu_baseline_reference(m, x, seconds)
    if ref is not None:
        return ref
    from oracle import oracle as orc
    threads = omp_threads()
    y = orc.ncsr(m, x, nthreads=threads)
    calls, t0 = 0, time.perf_counter()
    while True:
        orc.ncsr(m, x, nthreads=threads)
        calls += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    per = el / calls
    alg = 12.0 * m.nnz + 4.0 * (m.nrows + 1) + 8.0 * m.ncols + 8.0 * m.nrows
    return y, dict(value=round(2.0 * m.nnz / per / 1e9, 3), unit="GFLOP/s", cores=threads, kind="port",
                   sample=f"{calls} calls of ncsr (OpenMP static over rows) on the full matrix, {el:.1f} s",
                   ms_per_spmv=round(per * 1e3, 3), alg_gbs=round(alg / per / 1e9, 1), **cpu_info())


def sub_rows(m, rows):
    """The rows `rows` of m as a CSR matrix with m's columns."""
    rows = np.asarray(rows)
    lens = (m.row[rows + 1] - m.row[rows]).astype(np.int64)
    idx = np.concatenate([np.arange(m.row[r], m.row[r + 1]) for r in rows]) if rows.size else np.zeros(0, np.int64)
    row = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    return ax.CsrMatrix(row, m.col[idx], m.val[idx], ncols=m.ncols)


def errors(ref, y):
    """The kernels table's error columns: errAbs / errRel / rowInd as the
    reference's get_errors computes them (cudaSpmv.cu:474-495: it compares
    |y| with |y_ref|) and the signed norm-wise relative error the parity
    tests use (rel_err)."""
    if ref is None:
        return dict(rel_err=None)
    from oracle import oracle as orc
    a, r, pos = orc.get_errors(ref, y, sign_blind=True)
    return dict(rel_err=rel_err(ref, y), errAbs=a, errRel=r, rowInd=int(pos))


def summarize(info, nnz, sec, name, err, extra=None):
    gbs = info["alg_bytes"] / sec / 1e9
    r = dict(kernel=info["name"], label=name, us=round(sec * 1e6, 2), gflops=round(2.0 * nnz / sec / 1e9, 2),
             alg_gbs=round(gbs, 1), frac=round(gbs / PEAK_HBM_GBS, 4), beta=round(info["beta"], 4),
             stream_gbs=round(info["stream_bytes"] / sec / 1e9, 1), launches=info["launches"])
    r.update(err if isinstance(err, dict) else dict(rel_err=err))
    if extra:
        r.update(extra)
    return r


def measure(m, name, fmt, kw, dx, dy, iters, warmup, yref, mode=TIME_REGION, batch=1):
    """One row of the kernels table: plan, time (the plan's build time on the
    host, upload included, in `plan_s`), errors, and for the AX layout the
    per-SpMV x refresh as the reference's ordering time `ot`."""
    t0 = time.perf_counter()
    p = ax.Plan(m, fmt, **kw)
    plan_s = time.perf_counter() - t0
    info = p.info()
    p.spmv(dx, dy)
    L.call("axspmv_device_sync")
    err = errors(yref, dy.download(np.float64, m.nrows))
    p.time(dx, dy, iters=max(warmup, 1), warmup=0, mode=mode, batch=batch)
    t = p.time(dx, dy, iters=iters, warmup=0, mode=mode, batch=batch)
    extra = dict(plan_s=round(plan_s, 3), ot_us=0.0, device_mb=round(info["device_bytes"] / 1e6, 2))
    if kw.get("flags", 0) & ax.PLAN_AX_LAYOUT:
        extra["ot_us"] = round(p.time_op(L.OP_XREFRESH, iters=iters, warmup=2, x=dx).mean_ms * 1e3, 2)
    return p, info, summarize(info, m.nnz, t.mean_ms * 1e-3, name, err, extra)


def formats_table(m, x, dx, devbuild=True):
    """The reference's formats table (cudaSpmv.cu:2420-2436) for this matrix:
    each reference format's name, memory footprint mfp [MB], occupancy beta
    and conversion time -- the host builders' `ct` (OpenMP threads of this
    job) as the reference measures it, and the device builders' time where
    there is one."""
    rows = [dict(format="fcsr", mfp_mb=round((12.0 * m.nnz + 4.0 * (m.nrows + 1) + 8.0 * m.nrows) / 1e6, 2),
                 beta=1.0, ct_s=0.0)]

    def add(fd, dev=None, label=None):
        r = dict(format=label or fd["name"], mfp_mb=round(fd["mfp"] / 1e6, 2), beta=round(fd["beta"], 4),
                 ct_s=round(fd["ct"], 6))
        if dev is not None:
            r["device_ct_s"] = round(dev["ms"] * 1e-3, 6)
        rows.append(r)

    for chunk, reorder in ((64, True), (64, False)):
        k = ax.build_k1(m, chunk, reorder)
        add(k["fd"], label=f"{k['fd']['name']}_c{chunk}")
        del k
    for hbs in (4, 64):
        h = ax.build_axc(m, x, hbs, flags=L.BUILD_AX)
        d = ax.dev_build_axc(m, x, hbs) if devbuild else None
        add(h["fd"], d, label=f"faxc_hbs{hbs}")
        del h, d
    for bs, thw, th, mode in ((256, 16, 1, ax.AXT_UNC), (256, 8, 1, ax.AXT_UNC), (256, 4, 1, ax.AXT_UNC),
                              (256, 64, 4, ax.AXT_UNC), (256, 64, 8, ax.AXT_UNC), (256, 64, 1, ax.AXT_COM),
                              (256, 64, 4, ax.AXT_COM), (256, 64, 8, ax.AXT_COM)):
        h = ax.build_axt(m, x, bs, thw, th, mode, flags=L.BUILD_AX)
        d = ax.dev_build_axt(m, x, bs, thw, th, mode) if devbuild else None
        add(h["fd"], d)
        del h, d
    return rows


def extras_same_matrix(m, x, dx, yref, plain_us):
    """Forward + adjoint, SpMM and FP32 values on the same matrix, against
    plain FP64 SpMV (plain_us) of the flagship AXT kernel."""
    from oracle import oracle as orc
    out = {}
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    # forward + adjoint of 0.5*||Ax - b||^2
    b = ax.uniform(4, n, stream=ax.VEC_STREAM + 1)
    p = ax.Plan(m, ax.FMT_AXT, thw=4, flags=ax.PLAN_TRANSPOSE)
    db = ax.DeviceBuffer.from_array(b)
    bufs = dict(y=ax.DeviceBuffer(8 * n), r=ax.DeviceBuffer(8 * n), xbar=ax.DeviceBuffer(8 * nc),
                vbar=ax.DeviceBuffer(8 * nnz))
    fa = {}
    for label, op in (("fused", L.OP_FWD_ADJ), ("split", L.OP_FWD_ADJ_SPLIT)):
        t = p.time_op(op, iters=30, warmup=3, x=dx, b=db, **bufs)
        fa[label] = round(t.mean_ms * 1e3, 2)
    r = bufs["r"].download(np.float64, n)
    xbar = bufs["xbar"].download(np.float64, nc)
    r_ref = yref - b if yref is not None else None
    t_adj = p.time_op(L.OP_ADJOINT, iters=30, warmup=3, ybar=bufs["r"], xbar=bufs["xbar"])
    out["fwd_adj"] = dict(plain_spmv_us=plain_us, fused_us=fa["fused"], split_us=fa["split"],
                          adjoint_only_us=round(t_adj.mean_ms * 1e3, 2),
                          fused_over_plain=round(fa["fused"] / plain_us, 3),
                          r_rel_err=rel_err(r_ref, r) if r_ref is not None else None,
                          xbar_rel_err=rel_err(orc.ncsr_t(m, r_ref, nc), xbar) if r_ref is not None else None)
    log(f"fwd_adj: {out['fwd_adj']}")
    del bufs, p
    out["fwd_adj_rct"] = fwd_adj_rct(m, dx, db, yref, b)
    log(f"fwd_adj_rct: {out['fwd_adj_rct']}")
    # SpMM, FP64 and FP32 values
    sp = []
    for k in (8, 32, 64):
        X = ax.uniform(3, nc * k).reshape(nc, k)
        dX = ax.DeviceBuffer.from_array(X)
        dY = ax.DeviceBuffer(8 * n * k)
        rows = np.arange(0, n, max(1, n // 2000))
        Yref = orc.spmm(sub_rows(m, rows), X, nthreads=omp_threads())  # FP64 oracle on the sampled rows
        for vt in (ax.F64, ax.F32):
            # the library's SpMM candidates (AXT k-tiles, RCT tiles), in this value type
            cands = ax.autotune_defaults(L.OP_SPMM)
            for o in cands:
                o.vtype = vt
            best, times = ax.autotune(m, op=L.OP_SPMM, k=k, candidates=cands, iters=5)
            c = cands[best]
            pm = ax.Plan(m, c.format, **{f: getattr(c, f) for f in ("vtype", "ktile", "thw", "th", "tile_rows",
                                                                     "slices")})
            kt = c.ktile
            t = pm.time_op(L.OP_SPMM, iters=10, warmup=2, k=k, X=dX, Y=dY)
            Y = dY.download(np.float64, n * k).reshape(n, k)
            # accuracy on the sampled rows against the FP64 oracle (for FP32
            # values: the error of storing the values in FP32)
            err = rel_err(Yref, Y[rows])
            alg = (12 if vt == ax.F64 else 8) * nnz + 4 * (n + 1) + 16 * n * k
            us = t.mean_ms * 1e3
            sp.append(dict(k=k, vtype="f64" if vt == ax.F64 else "f32val", kernel=pm.info()["name"], ktile=kt,
                           rows_checked=int(rows.size),
                           us=round(us, 1),
                           gflops=round(2.0 * nnz * k / (us * 1e-6) / 1e9, 1),
                           alg_gbs=round(alg / (us * 1e-6) / 1e9, 1),
                           per_rhs_vs_plain=round(us / k / plain_us, 3), rel_err_vs_f64_oracle=err))
            log(f"spmm: {sp[-1]}")
        del dX, dY
    out["spmm"] = sp
    return out


def conversion(m, x, dx):
    """Format conversion time (the reference's `ct`): host builders against
    the device builders, and the per-SpMV x refresh (`ot`) of the AX layout."""
    out = []
    for label, host, dev in (
            ("faxtuh1hw4", lambda: ax.build_axt(m, x, 256, 4, 1, ax.AXT_UNC),
             lambda: ax.dev_build_axt(m, x, 256, 4, 1, ax.AXT_UNC)),
            ("faxtuh4hw64", lambda: ax.build_axt(m, x, 256, 64, 4, ax.AXT_UNC),
             lambda: ax.dev_build_axt(m, x, 256, 64, 4, ax.AXT_UNC)),
            ("faxtch1hw64bs256", lambda: ax.build_axt(m, x, 256, 64, 1, ax.AXT_COM),
             lambda: ax.dev_build_axt(m, x, 256, 64, 1, ax.AXT_COM)),
            ("faxc_hbs4", lambda: ax.build_axc(m, x, 4), lambda: ax.dev_build_axc(m, x, 4))):
        h = host()
        dev()  # first call pays for code loading
        d = dev()
        same = bool(np.array_equal(h["ax"], d["ax"]))
        out.append(dict(format=label, host_ms=round(h["fd"]["ct"] * 1e3, 2), host_threads=omp_threads(),
                        device_ms=round(d["ms"], 3), ax_identical=same))
        log(f"conversion: {out[-1]}")
        del h, d
    p = ax.Plan(m, ax.FMT_AXT, thw=4, flags=ax.PLAN_AX_LAYOUT)
    t = p.time_op(L.OP_XREFRESH, iters=30, warmup=3, x=dx)
    out.append(dict(format="axt_unc_h1_thw4_ax", x_refresh_us=round(t.mean_ms * 1e3, 2)))
    log(f"conversion: {out[-1]}")
    return out


def copy_bandwidth(reps: int = 20) -> dict:
    """A 256 MiB device-to-device copy (copy_kernel), host-clocked over reps
    back-to-back launches: the read + write bandwidth this box reaches."""
    n = 32 << 20  # 256 MiB of doubles
    src, dst = ax.DeviceBuffer(8 * n), ax.DeviceBuffer(8 * n)
    for _ in range(3):
        L.call("axspmv_copy_f64", src.ptr, dst.ptr, n, None)
    L.call("axspmv_device_sync")
    t0 = time.perf_counter()
    for _ in range(reps):
        L.call("axspmv_copy_f64", src.ptr, dst.ptr, n, None)
    L.call("axspmv_device_sync")
    dt = (time.perf_counter() - t0) / reps
    return dict(us=round(dt * 1e6, 1), gbs_read_plus_write=round(16.0 * n / dt / 1e9, 1))


def calibrate():
    """Copy bandwidth, and the flagship AXT kernel on the power-law structure
    with columns folded into a smaller x (footprint sweep): the gather bound."""
    res = {"copy_256MB": copy_bandwidth()}
    m = ax.powerlaw(1 << 20, 2.0, 2048, 20.0, seed=1)
    sweep = []
    for fold in (1, 2, 4, 8, 64):
        ncols = m.ncols // fold
        # same rows and row lengths, columns folded into [0, ncols / fold)
        mm = ax.CsrMatrix(m.row, (m.col % ncols).astype(np.uint32), m.val, ncols=ncols)
        x = ax.uniform(1, m.ncols)
        dx, dy = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer(8 * m.nrows)
        for name, fmt, kw in (FLAGSHIP_AXT, FLAGSHIP):
            p = ax.Plan(mm, fmt, **kw)
            p.time(dx, dy, iters=5, warmup=0, mode=TIME_REGION)
            t = p.time(dx, dy, iters=30, warmup=0, mode=TIME_REGION)
            sweep.append(dict(x_footprint_mb=round(8 * ncols / 2 ** 20, 3), kernel=p.info()["name"],
                              us=round(t.mean_ms * 1e3, 1),
                              ggathers=round(m.nnz / (t.mean_ms * 1e-3) / 1e9, 1)))
            log(f"calibrate: {sweep[-1]}")
    res["x_footprint_sweep"] = sweep
    return res


def pmc_probe(matrix: str = "powerlaw"):
    """A 256 MB copy (known bytes: the counter scale), then, on the 2^20-row
    power-law matrix (or the permuted 27-point stencil on 128^3), 5 SpMVs of
    each flagship plan (AXT UNC th = 1 in one layout and in 4 column slices,
    and RCT); with "spmm", 3 SpMMs at k = 8, 32 and 64 of the plans the
    autotuner picks on the power-law matrix; with "stencil_rcm", 5 SpMVs of
    the reordered RCT (r8192, one slice) and AXT UNC th = 1 plans on the
    permuted stencil; with "batch256", 5 SpMVs of the 256 banded matrices'
    block-diagonal RCT plan with its default 16-bit keys, then 5 with 32-bit
    keys; with "fwd_adj_rcm", 5 fused forward + adjoint passes of the
    reordered RCT plan on the stencil structure with values N(1, 0.1)."""
    n = 32 << 20
    src, dst = ax.DeviceBuffer(8 * n), ax.DeviceBuffer(8 * n)
    L.call("axspmv_copy_f64", src.ptr, dst.ptr, n, None)
    L.call("axspmv_device_sync")
    del src, dst
    if matrix == "fwd_adj_rcm":
        m3, _ = W.config3()
        m, x, b = W.config5(m3)
        del m3
        p = ax.Plan(m, ax.FMT_RCT, flags=ax.PLAN_TRANSPOSE | ax.PLAN_REORDER)
        dx, db = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer.from_array(b)
        outs = [ax.DeviceBuffer(8 * m.nrows) for _ in range(3)] + [ax.DeviceBuffer(8 * m.nnz)]
        for _ in range(5):
            p.fwd_adj(dx, db, *outs, fused=True)
        L.call("axspmv_device_sync")
        log(f"pmc-probe: fwd_adj {p.info()}")
        return
    if matrix == "batch256":
        mats = W.config4b()
        for kw in ({}, dict(flags=ax.PLAN_RCT_WIDE_KEY)):
            bt = ax.Batch(mats, fmt=ax.FMT_RCT, **kw)
            dxb = ax.DeviceBuffer.from_array(ax.uniform(6, int(bt.xoff[-1])))
            dyb = ax.DeviceBuffer(8 * int(bt.yoff[-1]))
            for _ in range(5):
                bt.spmv(dxb, dyb)
            L.call("axspmv_device_sync")
            log(f"pmc-probe: batch256 {kw} {bt.info()}")
            del bt, dxb, dyb
        return
    if matrix in ("stencil", "stencil_rcm"):
        m, x = W.config3()
    else:
        m, x = W.config2()
    dx, dy = ax.DeviceBuffer.from_array(x), ax.DeviceBuffer(8 * m.nrows)
    if matrix == "spmm":
        for k, kw in ((8, dict(tile_rows=1024, slices=4, ktile=8)), (32, dict(tile_rows=256, slices=1, ktile=16)),
                      (64, dict(tile_rows=256, slices=1, ktile=16))):
            p = ax.Plan(m, ax.FMT_RCT, **kw)
            dX = ax.DeviceBuffer.from_array(W.config4a_rhs(m.ncols, k))
            dY = ax.DeviceBuffer(8 * m.nrows * k)
            for _ in range(3):
                p.spmm(k, dX, dY)
            L.call("axspmv_device_sync")
            log(f"pmc-probe: spmm k={k} {p.info()['name']} ktile={kw['ktile']}")
            del p, dX, dY
        return
    plans = (FLAGSHIP_AXT, FLAGSHIP_AXT_SLICED, FLAGSHIP)
    if matrix == "stencil_rcm":
        plans = [r for r in REORDERED if r[0] in ("rct_r8192_s1_rcm", "axt_unc_h1_thw4_rcm")]
    for _, fmt, kw in plans:
        p = ax.Plan(m, fmt, **kw)
        for _ in range(5):
            p.spmv(dx, dy)
        L.call("axspmv_device_sync")
        log(f"pmc-probe: {p.info()}")


def suite_c1(tables, extras):
    """Config 1: the 2D Laplacian 256^2 -- below launch latency, so every
    kernel is timed as graph replays of 50 SpMVs."""
    from oracle import oracle as orc
    m1, x1 = W.config1()
    y1 = orc.ncsr(m1, x1)
    d1, e1 = ax.DeviceBuffer.from_array(x1), ax.DeviceBuffer(8 * m1.nrows)
    rows = []
    for name, fmt, kw in (("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
                          ("csr_vector", ax.FMT_CSR, dict(kernel=ax.KER_STREAM)),
                          ("axt_unc_h1_thw4", ax.FMT_AXT, dict(thw=4)), ("axt_unc_h4_thw64", ax.FMT_AXT, dict(thw=64, th=4)),
                          ("axt_unc_h1_thw4_ax", ax.FMT_AXT, dict(thw=4, flags=ax.PLAN_AX_LAYOUT)),
                          ("axc_hbs4", ax.FMT_AXC, dict(hbs=4)), ("k1_c64", ax.FMT_K1, dict(chunk=64)),
                          ("k1_c64_natural", ax.FMT_K1, dict(chunk=64, reorder=0)),
                          ("axt_com_h1_thw64_bs256", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
                          ("axt_com_h4_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4)),
                          ("rct_auto", ax.FMT_RCT, dict())):
        _, _, r = measure(m1, name, fmt, kw, d1, e1, 20, 2, y1, mode=TIME_GRAPH, batch=50)
        rows.append(r)
        log(f"laplace256: {r}")
    tables["laplace256"] = dict(config="C1: 2D 5-point Laplacian 256^2, x ~ U[-1,1] seed 0; graph-timed (50 SpMVs "
                                       "per replay)", nrows=m1.nrows, nnz=m1.nnz,
                                formats=formats_table(m1, x1, d1), kernels=rows)
    extras["laplace256_rocsparse"] = rocsparse_rows(m1, d1, y1, iters=200)


def suite_c3_c5(tables, extras):
    """Config 3 (the permuted 27-point stencil on 128^3, every format) and
    config 5 (forward + adjoint on its structure with values N(1, 0.1))."""
    from oracle import oracle as orc
    m3, x3 = W.config3()
    y3 = orc.ncsr(m3, x3, nthreads=omp_threads())
    d3, e3 = ax.DeviceBuffer.from_array(x3), ax.DeviceBuffer(8 * m3.nrows)
    rows = []
    c3 = [o for o in OTHERS if not o[0].startswith("rct_r") and not o[0].endswith("f32val")]
    for name, fmt, kw in [FLAGSHIP] + c3 + REORDERED:
        _, _, r = measure(m3, name, fmt, kw, d3, e3, 20, 3, y3)
        rows.append(r)
        log(f"stencil128p: {r}")
    tables["stencil128p"] = dict(config="C3: 27-point stencil 128^3, symmetric random permutation (seed 2), "
                                        "x ~ U[-1,1] seed 2", nrows=m3.nrows, nnz=m3.nnz,
                                 formats=formats_table(m3, x3, d3), kernels=rows)
    extras["stencil128p_rocsparse"] = rocsparse_rows(m3, d3, y3, iters=20)
    extras["stencil128p_spmm_k8"] = spmm_reordered(m3)
    del e3
    # config 5: forward + adjoint on the stencil structure, values N(1, 0.1)
    m5, x5, b5 = W.config5(m3)
    del m3
    y5 = orc.ncsr(m5, x5, nthreads=omp_threads())
    d5 = ax.DeviceBuffer.from_array(x5)
    e5 = ax.DeviceBuffer(8 * m5.nrows)
    for suffix, flags in (("", 0), ("_rcm", ax.PLAN_REORDER)):
        pl = ax.Plan(m5, ax.FMT_AXT, thw=4, flags=flags)
        pl.time(d5, e5, iters=3, warmup=0, mode=TIME_REGION)
        plain = pl.time(d5, e5, iters=20, warmup=0, mode=TIME_REGION).mean_ms * 1e3
        del pl
        extras["stencil128p_fwd_adj" + suffix] = fwd_adj_only(m5, x5, d5, y5, plain, b5, flags=flags)
        extras["stencil128p_fwd_adj_rct" + suffix] = fwd_adj_rct(m5, d5, ax.DeviceBuffer.from_array(b5), y5, b5,
                                                                 flags=flags)
        log(f"stencil fwd_adj_rct{suffix}: {extras['stencil128p_fwd_adj_rct' + suffix]}")


def spmm_reordered(m, k=8):
    """SpMM with k right-hand sides on the permuted stencil, plain and with
    AXSPMV_PLAN_REORDER (X's rows moved into the renumbered order and Y's
    back inside the call), against the FP64 oracle on sampled rows."""
    from oracle import oracle as orc
    n, nnz = m.nrows, m.nnz
    X = ax.uniform(3, m.ncols * k).reshape(m.ncols, k)
    dX, dY = ax.DeviceBuffer.from_array(X), ax.DeviceBuffer(8 * n * k)
    rows = np.arange(0, n, max(1, n // 2000))
    Yref = orc.spmm(sub_rows(m, rows), X, nthreads=omp_threads())
    out = []
    for name, fmt, kw in (("rct_r1024", ax.FMT_RCT, dict(tile_rows=1024, ktile=8)),
                          ("axt_unc_h1_thw4", ax.FMT_AXT, dict(thw=4, ktile=8))):
        for suffix, flags in (("", 0), ("_rcm", ax.PLAN_REORDER)):
            p = ax.Plan(m, fmt, flags=flags, **kw)
            t = p.time_op(L.OP_SPMM, iters=10, warmup=2, k=k, X=dX, Y=dY)
            Y = dY.download(np.float64, n * k).reshape(n, k)
            us = t.mean_ms * 1e3
            out.append(dict(k=k, label=name + suffix, kernel=p.info()["name"], us=round(us, 1),
                            gflops=round(2.0 * nnz * k / (us * 1e-6) / 1e9, 1),
                            rel_err_vs_f64_oracle=rel_err(Yref, Y[rows]), rows_checked=int(rows.size)))
            log(f"stencil spmm: {out[-1]}")
            del p
    return out


def suite_c4b(extras):
    """Config 4B: 256 uneven banded matrices in one launch -- the AXT
    persistent batch with uniform tiles, with the autotuner's per-group tiles
    (or its RCT choice), and the RCT block-diagonal plan."""
    from oracle import oracle as orc
    mats = W.config4b()
    t0 = time.perf_counter()
    tuned, ngroups = ax.autotune_batch(mats, iters=5)
    tune_s = round(time.perf_counter() - t0, 1)
    out = {}
    for label, make in (("uniform_thw4", lambda: ax.Batch(mats, thw=4)),
                        ("autotuned", lambda: ax.Batch(mats, opts_list=tuned)),
                        ("rct_block_diagonal", lambda: ax.Batch(mats, fmt=ax.FMT_RCT))):
        t0 = time.perf_counter()
        bt = make()
        build_s = round(time.perf_counter() - t0, 2)
        info = bt.info()
        xb = ax.uniform(6, int(bt.xoff[-1]))
        dxb, dyb = ax.DeviceBuffer.from_array(xb), ax.DeviceBuffer(8 * int(bt.yoff[-1]))
        bt.spmv(dxb, dyb)
        L.call("axspmv_device_sync")
        yb = dyb.download(np.float64, int(bt.yoff[-1]))
        errs = [rel_err(orc.ncsr(mats[i], xb[bt.xoff[i]:bt.xoff[i + 1]]), yb[bt.yoff[i]:bt.yoff[i + 1]])
                for i in range(0, 256, 16)]
        t = bt.time(dxb, dyb, iters=10, warmup=2, mode=TIME_REGION)
        tg = bt.time(dxb, dyb, iters=5, warmup=1, mode=TIME_GRAPH, batch=10)
        sec = t.mean_ms * 1e-3
        out[label] = dict(matrices=256, nnz=info["nnz"], items=info["nitems"], waves=info["nwaves"],
                          us=round(sec * 1e6, 1), graph_us=round(tg.mean_ms * 1e3, 1),
                          gflops=round(2.0 * info["nnz"] / sec / 1e9, 1),
                          alg_gbs=round(info["alg_bytes"] / sec / 1e9, 1),
                          frac=round(info["alg_bytes"] / sec / 1e9 / PEAK_HBM_GBS, 4),
                          max_rel_err_sampled=max(errs), members_checked=len(errs), build_s=build_s)
        if label == "autotuned":
            out[label].update(
                tuner_option_sets=ngroups, tuner_s=tune_s,
                chosen="rct" if tuned[0].format == ax.FMT_RCT else
                f"axt per group, (thw, th) used: {sorted({(o.thw, o.th) for o in tuned})}")
        log(f"batch256 {label}: {out[label]}")
        del dxb, dyb, bt
    extras["batch256"] = out


def fwd_adj_rct(m, dx, db, yref, b, flags=0):
    """Forward + adjoint on RCT plans (the fused form: SpMV, residual, one
    sweep over the RCT plan of A^T adding x_bar and writing v_bar; the split
    form: SpMV, residual, adjoint, value gradient), against plain RCT."""
    from oracle import oracle as orc
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    p = ax.Plan(m, ax.FMT_RCT, flags=ax.PLAN_TRANSPOSE | flags)
    dy = ax.DeviceBuffer(8 * n)
    p.time(dx, dy, iters=3, warmup=0, mode=TIME_REGION)
    plain = p.time(dx, dy, iters=20, warmup=0, mode=TIME_REGION).mean_ms * 1e3
    bufs = dict(y=dy, r=ax.DeviceBuffer(8 * n), xbar=ax.DeviceBuffer(8 * nc), vbar=ax.DeviceBuffer(8 * nnz))
    split = p.time_op(L.OP_FWD_ADJ_SPLIT, iters=20, warmup=2, x=dx, b=db, **bufs)
    t = p.time_op(L.OP_FWD_ADJ, iters=20, warmup=2, x=dx, b=db, **bufs)
    out = dict(plan=p.info()["name"], plain_spmv_us=round(plain, 2), fwd_adj_us=round(t.mean_ms * 1e3, 2),
               split_us=round(split.mean_ms * 1e3, 2), fwd_adj_over_plain=round(t.mean_ms * 1e3 / plain, 3))
    if yref is not None:
        r_ref = yref - b
        out["r_rel_err"] = rel_err(r_ref, bufs["r"].download(np.float64, n))
        out["xbar_rel_err"] = rel_err(orc.ncsr_t(m, r_ref, nc), bufs["xbar"].download(np.float64, nc))
    return out


def fwd_adj_only(m, x, dx, yref, plain_us, b=None, flags=0):
    """Forward + adjoint on the AXT UNC th = 1 plan: fused (y, r and v_bar in
    one sweep, then x_bar = A^T r) and split (four kernels)."""
    from oracle import oracle as orc
    n, nc, nnz = m.nrows, m.ncols, m.nnz
    if b is None:
        b = ax.uniform(4, n, stream=ax.VEC_STREAM + 1)
    p = ax.Plan(m, ax.FMT_AXT, thw=4, flags=ax.PLAN_TRANSPOSE | flags)
    db = ax.DeviceBuffer.from_array(b)
    bufs = dict(y=ax.DeviceBuffer(8 * n), r=ax.DeviceBuffer(8 * n), xbar=ax.DeviceBuffer(8 * nc),
                vbar=ax.DeviceBuffer(8 * nnz))
    fa = {}
    for label, op in (("fused", L.OP_FWD_ADJ), ("split", L.OP_FWD_ADJ_SPLIT)):
        t = p.time_op(op, iters=20, warmup=2, x=dx, b=db, **bufs)
        fa[label] = round(t.mean_ms * 1e3, 2)
    r = bufs["r"].download(np.float64, n)
    xbar = bufs["xbar"].download(np.float64, nc)
    r_ref = yref - b
    out = dict(plan=p.info()["name"], plain_spmv_us=round(plain_us, 2), fused_us=fa["fused"], split_us=fa["split"],
               fused_over_plain=round(fa["fused"] / plain_us, 3), r_rel_err=rel_err(r_ref, r),
               xbar_rel_err=rel_err(orc.ncsr_t(m, r_ref, nc), xbar))
    log(f"stencil fwd_adj: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--formats", choices=["flagship", "all"], default="all")
    ap.add_argument("--suite", choices=["all", "headline"], default="all")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--pmc-probe", nargs="?", const="powerlaw", choices=["powerlaw", "stencil", "spmm", "stencil_rcm",
                                                                           "batch256", "fwd_adj_rcm"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: SpMV here runs on a single MI355X; "
              "multi-GPU row-block sharding is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate:
        print(json.dumps(calibrate()), flush=True)
        return
    if args.pmc_probe:
        pmc_probe(args.pmc_probe)
        return

    t_start = time.perf_counter()
    copy = copy_bandwidth()
    log(f"copy: {copy}")
    m, x = W.config2()
    log(f"matrix: n={m.nrows} nnz={m.nnz} ({time.perf_counter() - t_start:.1f} s)")

    yref, cpu = None, None
    if not args.no_cpu:
        yref, cpu = cpu_baseline(m, x, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    dx = ax.DeviceBuffer.from_array(x)
    dy = ax.DeviceBuffer(8 * m.nrows)
    name, fmt, kw = FLAGSHIP
    plan, info, flag = measure(m, name, fmt, kw, dx, dy, 100, args.warmup, yref)
    log(f"flagship: {flag}")

    # The timed region: K SpMVs back to back, device-synchronized on both
    # sides.  HIP events on the launch stream mark the start of every SpMV
    # and, for a sliced RCT plan (the tile kernel, then the merge of its slice
    # partials), the end of its tile kernel, so the whole SpMV's mean, its
    # spread and the tile kernel's own mean come out of the same region
    # (TIME_KERNEL).
    plan.time(dx, dy, iters=max(args.warmup, 1), warmup=0, mode=TIME_REGION)
    L.call("axspmv_device_sync")
    w0 = time.perf_counter()
    t = plan.time(dx, dy, iters=args.steps, warmup=0, mode=TIME_KERNEL)
    L.call("axspmv_device_sync")
    wall = time.perf_counter() - w0
    kern_s = t.main_ms * 1e-3
    path_s = t.mean_ms * 1e-3
    achieved = info["alg_bytes"] / path_s / 1e9
    gflops = 2.0 * m.nnz / (wall / args.steps) / 1e9
    # the same SpMV with the Infinity Cache and L2 flushed before each launch
    tc = plan.time(dx, dy, iters=20, warmup=0, mode=TIME_COLD)
    cold = dict(spmv_us_mean=round(tc.mean_ms * 1e3, 2), spmv_us_median=round(tc.median_ms * 1e3, 2),
                alg_gbs=round(info["alg_bytes"] / (tc.median_ms * 1e-3) / 1e9, 1),
                frac=round(info["alg_bytes"] / (tc.median_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4))
    log(f"cold: {cold}")
    del plan

    formats = [flag]
    extras, tables = {}, {}
    if args.formats == "all":
        for (n2, f2, k2) in OTHERS:
            try:
                _, _, r = measure(m, n2, f2, k2, dx, dy, 50, 5, yref)
            except ax.AxspmvError as e:
                r = dict(label=n2, error=str(e))
            log(f"{n2}: {r}")
            formats.append(r)
        tables["powerlaw"] = dict(config="C2: power-law n = 2^20, Zipf(2) row lengths on {1..2048}, mean 20, "
                                         "x ~ U[-1,1] seed 1", nrows=m.nrows, nnz=m.nnz,
                                  formats=formats_table(m, x, dx), kernels=formats)
        plain = next(f["us"] for f in formats if f.get("label") == FLAGSHIP_AXT[0])
        extras.update(extras_same_matrix(m, x, dx, yref, plain))
        extras["conversion"] = conversion(m, x, dx)
        extras["rocsparse"] = rocsparse_rows(m, dx, yref)
    del dx, dy
    if args.suite == "all":
        del m
        suite_c1(tables, extras)
        suite_c3_c5(tables, extras)
        suite_c4b(extras)

    traffic, traffic_src, traffic_kernel = None, None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_powerlaw_flagship.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        entry = tj.get(info["name"], {})
        if "hbm_bytes_per_launch" in entry:
            # every launch of one SpMV: the tile kernel and, with slices, the
            # merge of the slice partials (or the y clear before atomic adds)
            other = entry.get("merge_kernel", entry.get("clear_kernel", {}))
            traffic = entry["hbm_bytes_per_launch"] + other.get("fetch_bytes", 0) + other.get("write_bytes", 0)
            traffic_kernel = entry.get("kernel")
            traffic_src = os.path.relpath(tpath, ROOT)

    us = lambda v: round(v * 1e3, 2)  # noqa: E731  (ms -> us)
    out = {
        "metric": "SpMV GFLOP/s + HBM GB/s (% of 8 TB/s) per format; fwd+adjoint vs plain",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded power-law matrix generated on the host)",
        "config": {"workload": "powerlaw n=2^20 Zipf(2) rl<=2048 mean 20 seed 1, FP64 SpMV",
                   "nrows": info["nrows"], "nnz": info["nnz"], "kernel": info["name"],
                   "alg_bytes_per_spmv": info["alg_bytes"], "parity_rel_err": flag["rel_err"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "scope": "whole SpMV (every launch of one SpMV: tile kernel + merge of the slice partials "
                              "for a sliced RCT plan)",
                     "spmv_us": round(path_s * 1e6, 2), "spmv_us_median": us(t.median_ms),
                     "spmv_us_min": us(t.min_ms), "spmv_us_max": us(t.max_ms),
                     "kernel": traffic_kernel or info["name"], "kernel_us": round(kern_s * 1e6, 2),
                     "kernel_frac": round(info["alg_bytes"] / kern_s / 1e9 / PEAK_HBM_GBS, 4),
                     "alg_bytes_per_spmv": info["alg_bytes"], "launches_per_spmv": info["launches"],
                     "cold_caches": cold, "traffic_source": traffic_src,
                     "copy_same_box": dict(copy, frac_of_copy=round(achieved / copy["gbs_read_plus_write"], 4))},
        "cpu_baseline": cpu,
        "formats": formats,
        "tables": tables,
        "extras": extras,
        "bench_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
