"""Sweep driver and tuning database (the counterpart of jobLauncher.sh).

The reference's launcher runs its binary once per matrix file and block
size (jobLauncher.sh:58-68 lists the matrices, 145-150 the block sizes 32 to
1024, 174 the call `./EXE BS MAT`); every run builds all formats and prints
one result row per kernel (cudaSpmv.cu:2420-2492).  Here one process does
the whole sweep on one GPU:

    python -m gpu_spmv_amd.sweep --out db.json M01.bin M02.csr ...
    python -m gpu_spmv_amd.sweep --out db.json --generate laplace2d:256 powerlaw:65536

For every matrix: its statistics (the `.sta` file of printMatrixStats,
written next to the database with --sta), then every candidate (format,
parameters, block size) as a plan, timed with one event pair around
back-to-back SpMVs; each result row holds the kernel name, time, GFLOP/s,
algorithmic GB/s and the error against the CPU product of the same matrix
(`check` selects that product; the default, ``"first"``, compares every
kernel with the first kernel that ran, so the sweep needs nothing but the
library).  The database maps each matrix to its rows and to the fastest
candidate, and `best_opts` reads it back.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib as L
from .core import (TIME_REGION, CsrMatrix, DeviceBuffer, Plan, banded, laplace2d, powerlaw, random_sparse,
                   stencil27, uniform)

BLOCKS = (32, 64, 128, 256, 512, 1024)

# (label, format, options, block sensitive); block-sensitive candidates are
# run once per block size they accept (RCT: multiples of 64), the others
# once (their kernels fix the workgroup).
CANDIDATES = [
    ("csr_naive", L.FMT_CSR, dict(kernel=L.KER_NAIVE), True),
    ("csr_vector", L.FMT_CSR, dict(kernel=L.KER_STREAM), False),
    ("k1_c64", L.FMT_K1, dict(chunk=64), False),
    ("axc_hbs4", L.FMT_AXC, dict(hbs=4), False),
    ("axc_hbs32", L.FMT_AXC, dict(hbs=32), False),
    ("axt_unc_h1_thw4", L.FMT_AXT, dict(thw=4, th=1), False),
    ("axt_unc_h1_thw8", L.FMT_AXT, dict(thw=8, th=1), False),
    ("axt_unc_h1_thw16", L.FMT_AXT, dict(thw=16, th=1), False),
    ("axt_unc_h4_thw64", L.FMT_AXT, dict(thw=64, th=4), False),
    ("axt_unc_h8_thw64", L.FMT_AXT, dict(thw=64, th=8), False),
    ("axt_unc_h1_thw2_s4", L.FMT_AXT, dict(thw=2, slices=4), False),
    ("axt_unc_h1_thw2_s8", L.FMT_AXT, dict(thw=2, slices=8), False),
    ("axt_com_h1_thw64_bs256", L.FMT_AXT, dict(mode=L.AXT_COM, thw=64, th=1, bs=256), False),
    ("axt_com_h4_thw64", L.FMT_AXT, dict(mode=L.AXT_COM, thw=64, th=4), False),
    ("rct_auto", L.FMT_RCT, dict(), True),
    # renumbered by reverse Cuthill-McKee inside the plan (square matrices)
    ("rct_auto_rcm", L.FMT_RCT, dict(flags=L.PLAN_REORDER), False),
    ("axt_unc_h1_thw4_rcm", L.FMT_AXT, dict(thw=4, flags=L.PLAN_REORDER), False),
]

GENERATORS = {
    "laplace2d": lambda a: laplace2d(int(a)),
    "powerlaw": lambda a: powerlaw(int(a), 2.0, 2048, 20.0, seed=1),
    "stencil27p": lambda a: stencil27(int(a), permute=True, seed=2),
    "banded": lambda a: banded(int(a), 32, 0.3, seed=5),
    "random": lambda a: random_sparse(int(a), 1, 64, seed=7),
}


def load_matrix(spec: str) -> tuple[str, CsrMatrix]:
    """A `.csr` / `.bin` path, or `generator:size` (see GENERATORS)."""
    if ":" in spec and not os.path.exists(spec):
        gen, arg = spec.split(":", 1)
        if gen not in GENERATORS:
            raise SystemExit(f"unknown generator {gen!r}: one of {sorted(GENERATORS)}")
        return f"{gen}_{arg}", GENERATORS[gen](arg)
    name = os.path.splitext(os.path.basename(spec))[0]
    return name, CsrMatrix.read(spec)


def _err(ref: np.ndarray, y: np.ndarray) -> float:
    den = float(np.max(np.abs(ref))) if ref.size else 0.0
    d = float(np.max(np.abs(y - ref))) if ref.size else 0.0
    return d / den if den else d


def sweep_matrix(name: str, m: CsrMatrix, blocks=BLOCKS, candidates=CANDIDATES, iters: int = 50,
                 warmup: int = 5, check=None, log=None) -> dict:
    """Time every candidate on one matrix.  check(m, x) -> y_ref, or None to
    compare with the first kernel's result."""
    x = uniform(1, m.ncols)
    dx, dy = DeviceBuffer.from_array(x), DeviceBuffer(8 * max(m.nrows, 1))
    yref = check(m, x) if check is not None else None
    rows = []
    for label, fmt, kw, by_block in candidates:
        ok_blocks = tuple(b for b in blocks if fmt != L.FMT_RCT or b % 64 == 0) or (0,)
        for bs in (ok_blocks if by_block else (0,)):
            opts = dict(kw)
            if bs:
                opts["block"] = bs
            row = dict(label=label, format=fmt, opts=opts)
            try:
                p = Plan(m, fmt, **opts)
            except L.AxspmvError as e:
                row["error"] = str(e)
                rows.append(row)
                continue
            info = p.info()
            p.spmv(dx, dy)
            L.call("axspmv_device_sync")
            y = dy.download(np.float64, m.nrows)
            if yref is None:
                yref = y.copy()
            p.time(dx, dy, iters=max(warmup, 1), warmup=0, mode=TIME_REGION)
            t = p.time(dx, dy, iters=iters, warmup=0, mode=TIME_REGION)
            sec = t.mean_ms * 1e-3
            row.update(kernel=info["name"], us=round(sec * 1e6, 3), gflops=round(2.0 * m.nnz / sec / 1e9, 3),
                       alg_gbs=round(info["alg_bytes"] / sec / 1e9, 2), beta=round(info["beta"], 4),
                       rel_err=_err(yref, y))
            rows.append(row)
            if log:
                log(f"{name}: {row}")
            del p
    ok = [r for r in rows if "us" in r]
    best = min(ok, key=lambda r: r["us"]) if ok else None
    return dict(nrows=m.nrows, ncols=m.ncols, nnz=m.nnz, rows=rows,
                best=dict(label=best["label"], format=best["format"], opts=best["opts"], us=best["us"])
                if best else None)


def best_opts(db: dict | str, matrix: str) -> tuple[int, dict] | None:
    """(format, options) of the fastest candidate recorded for `matrix`."""
    if isinstance(db, str):
        with open(db) as f:
            db = json.load(f)
    b = db.get("matrices", {}).get(matrix, {}).get("best")
    return (b["format"], dict(b["opts"])) if b else None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("matrices", nargs="*", help=".csr / .bin files")
    ap.add_argument("--generate", nargs="*", default=[], help="generator:size, e.g. laplace2d:256")
    ap.add_argument("--blocks", default=",".join(map(str, BLOCKS)))
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--out", required=True, help="tuning database (JSON)")
    ap.add_argument("--sta", action="store_true", help="also write <matrix>.sta next to the database")
    args = ap.parse_args(argv)
    blocks = tuple(int(b) for b in args.blocks.split(",") if b)
    specs = list(args.matrices) + list(args.generate)
    if not specs:
        ap.error("no matrices")
    db = dict(created=time.strftime("%Y-%m-%d %H:%M:%S"), blocks=list(blocks), matrices={})
    outdir = os.path.dirname(os.path.abspath(args.out))
    for spec in specs:
        name, m = load_matrix(spec)
        if args.sta:
            L.call("axspmv_csr_write_sta", m.ref, name.encode(), os.path.join(outdir, name + ".sta").encode())
        db["matrices"][name] = sweep_matrix(name, m, blocks, iters=args.iters,
                                            log=lambda s: print(s, file=sys.stderr, flush=True))
        with open(args.out, "w") as f:
            json.dump(db, f, indent=1)
        print(f"{name}: best {db['matrices'][name]['best']}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
