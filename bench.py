"""SpMV benchmark on one MI355X.

Workload (one step = one SpMV y = A x, inputs resident in HBM):
  synthetic power-law matrix, n = 2^20 rows, row lengths Zipf(alpha = 2) on
  {1..2048} rescaled to mean 20 (nnz ~ 21.0M), uniform sorted distinct
  columns, values N(0,1), x ~ U[-1,1], seed 1 -- FP64, in the AXT format
  (uncompacted, tile height 1, tile half width 4).

Prints one JSON line: GFLOP/s of the flagship kernel (2*nnz per SpMV), its
roofline against 8 TB/s of HBM (algorithmic bytes per SpMV / kernel time
from HIP events), the CPU baseline (the oracle's CSR product on this host's
cores), and every other format on the same matrix.

  python bench.py [--steps K] [--warmup W] [--formats flagship|all]
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
from gpu_spmv_amd.core import TIME_GRAPH, TIME_REGION  # noqa: E402

PEAK_HBM_GBS = 8000.0
FLAGSHIP = ("axt_unc_h1_thw4", ax.FMT_AXT, dict(thw=4, th=1))
OTHERS = [
    ("csr_naive", ax.FMT_CSR, dict(kernel=ax.KER_NAIVE)),
    ("axt_unc_h1_thw8", ax.FMT_AXT, dict(thw=8, th=1)),
    ("axt_unc_h1_thw16", ax.FMT_AXT, dict(thw=16, th=1)),
    ("axt_unc_h4_thw64", ax.FMT_AXT, dict(thw=64, th=4)),
    ("axt_unc_h8_thw64", ax.FMT_AXT, dict(thw=64, th=8)),
    ("axt_unc_h1_thw4_f32val", ax.FMT_AXT, dict(thw=4, th=1, vtype=ax.F32)),
    ("axt_com_h1_thw64_bs256", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=1, bs=256)),
    ("axt_com_h4_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=4)),
    ("axt_com_h8_thw64", ax.FMT_AXT, dict(mode=ax.AXT_COM, thw=64, th=8)),
    ("k1_c64", ax.FMT_K1, dict(chunk=64)),
    ("k1_c64_natural", ax.FMT_K1, dict(chunk=64, reorder=0)),
    ("axc_hbs4", ax.FMT_AXC, dict(hbs=4)),
    ("axc_hbs64", ax.FMT_AXC, dict(hbs=64)),
    ("rct_auto", ax.FMT_RCT, dict()),
    ("rct_r16384_s4", ax.FMT_RCT, dict(tile_rows=16384, slices=4)),
    ("rct_r8192_s4", ax.FMT_RCT, dict(tile_rows=8192, slices=4)),
    ("rct_r8192_s8", ax.FMT_RCT, dict(tile_rows=8192, slices=8)),
    ("rct_r16384_s2", ax.FMT_RCT, dict(tile_rows=16384, slices=2)),
    ("rct_r4096_s1", ax.FMT_RCT, dict(tile_rows=4096, slices=1)),
]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err(ref, y):
    return float(np.max(np.abs(y - ref)) / np.max(np.abs(ref)))


def cpu_baseline(m, x, seconds: float):
    """The oracle's CSR product (the reference's ncsr) on all host threads."""
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1)
    y = orc.ncsr(m, x, nthreads=threads)
    calls, t0 = 0, time.perf_counter()
    while True:
        orc.ncsr(m, x, nthreads=threads)
        calls += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    per = el / calls
    return y, dict(value=round(2.0 * m.nnz / per / 1e9, 3), unit="GFLOP/s", cores=threads, kind="port",
                   sample=f"{calls} calls of ncsr (OpenMP static over rows) on the full matrix, {el:.1f} s",
                   ms_per_spmv=round(per * 1e3, 3))


def measure(m, name, fmt, kw, dx, dy, iters, warmup, yref):
    p = ax.Plan(m, fmt, **kw)
    info = p.info()
    p.spmv(dx, dy)
    ax._lib.call("axspmv_device_sync")
    err = rel_err(yref, dy.download(np.float64, m.nrows)) if yref is not None else None
    p.time(dx, dy, iters=max(warmup, 1), warmup=0, mode=TIME_REGION)
    t = p.time(dx, dy, iters=iters, warmup=0, mode=TIME_REGION)
    sec = t.mean_ms * 1e-3
    gbs = info["alg_bytes"] / sec / 1e9
    return p, info, dict(kernel=info["name"], label=name, us=round(t.mean_ms * 1e3, 2),
                         gflops=round(2.0 * m.nnz / sec / 1e9, 2), alg_gbs=round(gbs, 1),
                         frac=round(gbs / PEAK_HBM_GBS, 4), beta=round(info["beta"], 4),
                         stream_gbs=round(info["stream_bytes"] / sec / 1e9, 1),
                         rel_err=err)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--formats", choices=["flagship", "all"], default="all")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: SpMV here runs on a single MI355X; "
              "multi-GPU row-block sharding is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    m = ax.powerlaw(1 << 20, 2.0, 2048, 20.0, seed=1)
    x = ax.uniform(1, m.ncols)
    log(f"matrix: n={m.nrows} nnz={m.nnz} ({time.perf_counter() - t0:.1f} s)")

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
    # sides; HIP events on the launch stream bracket the same region.
    plan.time(dx, dy, iters=max(args.warmup, 1), warmup=0, mode=TIME_REGION)
    ax._lib.call("axspmv_device_sync")
    w0 = time.perf_counter()
    t = plan.time(dx, dy, iters=args.steps, warmup=0, mode=TIME_REGION)
    ax._lib.call("axspmv_device_sync")
    wall = time.perf_counter() - w0
    kern_s = t.mean_ms * 1e-3
    achieved = info["alg_bytes"] / kern_s / 1e9
    gflops = 2.0 * m.nnz / (wall / args.steps) / 1e9

    formats = [flag]
    if args.formats == "all":
        for (n2, f2, k2) in OTHERS:
            try:
                _, _, r = measure(m, n2, f2, k2, dx, dy, 50, 5, yref)
            except ax.AxspmvError as e:
                r = dict(label=n2, error=str(e))
            log(f"{n2}: {r}")
            formats.append(r)

    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_axt_unc_h1_thw4_powerlaw.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic, traffic_src = tj.get("hbm_bytes_per_launch"), os.path.relpath(tpath, ROOT)

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
        "config": {"workload": "powerlaw n=2^20 Zipf(2) rl<=2048 mean 20 seed 1, FP64 SpMV, AXT UNC th=1 thw=4",
                   "nrows": m.nrows, "nnz": m.nnz, "kernel": info["name"],
                   "alg_bytes_per_spmv": info["alg_bytes"], "parity_rel_err": flag["rel_err"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel_us": round(t.mean_ms * 1e3, 2), "traffic_source": traffic_src},
        "cpu_baseline": cpu,
        "formats": formats,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
