"""Write the golden fixtures in tests/golden/*.npz from the reference itself.

Runs oracle/_ref/ref_dump_c32 and _c64 -- the host-only parts of the
reference's cudaSpmv.cu compiled by `make -C oracle ref` -- on small
matrices and stores, per matrix, the inputs (row/col/val, x) and the
reference's outputs: ncsr's y, rl[] and the .sta statistics, and the K1, AXC
and AXT arrays at the reference's 32-lane parameters and at the 64-lane
parameters the kernels use; and, in csrtext.npz, two .csr text files with
the arrays the reference's matrixReading parses from them.

Needs /root/reference (present only where the fixtures are made); the tests
read the committed .npz files and never run the reference.

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import gpu_spmv_amd as ax  # noqa: E402  (generators and .bin writer only)

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref")

AXC_HBS = [32, 64, 4, 8, 16]
UNC_H1 = [4, 8, 16, 32, 64]
UNC_H = [(32, 4), (32, 6), (32, 8), (32, 10), (32, 12), (64, 4), (64, 8)]
COM_H1 = [(32, 256), (32, 1024), (64, 256), (64, 64), (4, 32)]
COM_H = [(32, 4), (32, 8), (32, 16), (64, 4), (64, 8), (64, 16)]


def toy() -> ax.CsrMatrix:
    """12 rows: row 0 of length 1, empty rows, rows crossing tile boundaries."""
    rls = [1, 0, 5, 3, 9, 2, 0, 12, 1, 1, 4, 7]
    rng = np.random.default_rng(12)
    row = np.concatenate([[0], np.cumsum(rls)]).astype(np.uint32)
    col = np.concatenate([np.sort(rng.choice(12, k, replace=False)) for k in rls]).astype(np.uint32)
    val = rng.standard_normal(col.size)
    return ax.CsrMatrix(row, col, val, ncols=12)


def matrices():
    yield "toy12", toy(), "full"
    yield "rand200", ax.random_sparse(200, 0, 40, seed=7), "full"
    yield "lap16", ax.laplace2d(16), "full"
    yield "band300", ax.banded(300, 8, 0.3, seed=11), "full"
    yield "plaw512", ax.powerlaw(512, 2.0, 128, 10.0, seed=3), "some"


def run(binary, mat, xf, out, *args):
    subprocess.run([os.path.join(REF, binary), mat, xf, out, *map(str, args)], check=True,
                   stdout=subprocess.DEVNULL)


def read_meta(path):
    meta = {}
    with open(path) as f:
        for line in f:
            k, v = line.split(None, 1)
            v = v.strip()
            try:
                meta[k] = int(v)
            except ValueError:
                try:
                    meta[k] = float(v)
                except ValueError:
                    meta[k] = v
    return meta


def load_arrays(out, names):
    res = {}
    for n, dt in names:
        res[n] = np.fromfile(f"{out}.{n}", dtype=dt)
    return res


def varied_csr_text(seed: int = 31) -> str:
    """A .csr text file in the reference's layout ("nrows nnz", nnz lines
    "val col", nrows + 1 offsets) written with varied number formats and
    whitespace: 17-digit, exponent, signed, integer-valued, denormal and
    tiny values, tabs, blank lines, offsets spread over several lines."""
    rng = np.random.default_rng(seed)
    rls = [3, 0, 1, 5, 2, 0, 4, 6, 1, 3]
    row = np.concatenate([[0], np.cumsum(rls)])
    fmts = ["{:.17g}", "{:.6e}", "{:+.3f}", "{:g}", "{:.20e}", "{:.1f}", "{!r}"]
    vals = list(rng.standard_normal(sum(rls)) * 10.0 ** rng.integers(-300, 300, sum(rls)))
    vals[0], vals[1], vals[2], vals[3] = 5e-324, -0.0, 1.0, 2.5e-310
    lines = [f"  {len(rls)}\t{sum(rls)}"]
    k = 0
    for r, l in enumerate(rls):
        for c in sorted(rng.choice(40, l, replace=False)):
            sep = ["  ", "\t", " "][k % 3]
            lines.append(fmts[k % len(fmts)].format(float(vals[k])) + sep + str(int(c)) + ("\n" if k % 7 == 3 else ""))
            k += 1
    offs = [str(int(v)) for v in row]
    lines.append(" ".join(offs[:4]))
    lines.append("\t".join(offs[4:]))
    return "\n".join(lines) + "\n"


def csr_text_fixture(tmp):
    """tests/golden/csrtext.npz: .csr text files and the arrays the
    reference's matrixReading (cudaSpmv.cu:265-328) parses from them --
    one with varied formatting, one written by the library's .csr writer."""
    store = {}
    writer_m = ax.random_sparse(40, 0, 8, seed=21)
    path = os.path.join(tmp, "t.csr")
    writer_m.write(path)
    texts = {"varied": varied_csr_text(), "writer": open(path).read()}
    for name, text in texts.items():
        with open(path, "w") as f:
            f.write(text)
        out = os.path.join(tmp, "p")
        run("ref_dump_c32", path, "-", out, "parse")
        store[f"{name}.text"] = np.array(text)
        store.update({f"{name}.{k}": v for k, v in load_arrays(
            out, [("row", np.uint32), ("col", np.uint32), ("val", np.float64)]).items()})
    dst = os.path.join(HERE, "csrtext.npz")
    np.savez_compressed(dst, **store)
    print(f"csrtext: {', '.join(texts)} -> {os.path.getsize(dst)} bytes")


def main():
    if not (os.path.exists(os.path.join(REF, "ref_dump_c32")) and os.path.exists(os.path.join(REF, "ref_dump_c64"))):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    # Short, dot-free scratch path: the reference keeps file names in char[48]
    # and derives the .sta name by cutting at the first '.'.
    tmp = tempfile.mkdtemp(prefix="axref", dir="/tmp")
    assert "." not in tmp and len(tmp) < 30
    for name, m, scope in matrices():
        n = m.nrows
        x = ax.uniform(100 + n, n)
        mat, xf = os.path.join(tmp, "m.bin"), os.path.join(tmp, "x.bin")
        m.write(mat)
        x.tofile(xf)
        store = dict(row=m.row, col=m.col, val=m.val, x=x)
        metas = {}
        out = os.path.join(tmp, "o")
        run("ref_dump_c32", mat, xf, out, "ncsr")
        store.update({f"ncsr.{k}": v for k, v in load_arrays(out, [("y", np.float64), ("rl", np.uint32)]).items()})
        metas["ncsr"] = read_meta(out + ".meta")
        with open(os.path.join(tmp, "m.sta")) as f:
            store["ncsr.sta"] = np.array(f.read())
        for c in (32, 64):
            key = f"k1_c{c}"
            run(f"ref_dump_c{c}", mat, xf, out, "k1")
            store.update({f"{key}.{k}": v for k, v in load_arrays(
                out, [("permi", np.uint32), ("nmc", np.uint32), ("chp", np.uint32), ("val", np.float64),
                      ("col", np.uint32)]).items()})
            metas[key] = read_meta(out + ".meta")
        for hbs in (AXC_HBS if scope == "full" else [32, 64]):
            key = f"axc_h{hbs}"
            run("ref_dump_c32", mat, xf, out, "axc", hbs)
            store.update({f"{key}.{k}": v for k, v in load_arrays(
                out, [("ax", np.float64), ("brp", np.uint32), ("mapx", np.uint32)]).items()})
            metas[key] = read_meta(out + ".meta")
        reqs = [("UNC", 256, thw, 1) for thw in UNC_H1] + [("UNC", 256, thw, th) for thw, th in UNC_H] + \
               [("COM", bs, thw, 1) for thw, bs in COM_H1] + [("COM", 256, thw, th) for thw, th in COM_H]
        if scope != "full":
            reqs = [("UNC", 256, 4, 1), ("UNC", 256, 64, 4), ("COM", 256, 64, 1), ("COM", 256, 64, 8)]
        for mode, bs, thw, th in reqs:
            key = f"axt_{mode}_bs{bs}_hw{thw}_h{th}"
            run("ref_dump_c32", mat, xf, out, "axt", bs, thw, th, mode)
            store.update({f"{key}.{k}": v for k, v in load_arrays(
                out, [("ax", np.float64), ("sec", np.uint32), ("con", np.uint32)]).items()})
            metas[key] = read_meta(out + ".meta")
        store["meta"] = np.array(json.dumps(metas, sort_keys=True))
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **store)
        print(f"{name}: n={n} nnz={m.nnz} -> {os.path.getsize(path)} bytes")
    csr_text_fixture(tmp)
    for f in os.listdir(tmp):
        os.remove(os.path.join(tmp, f))
    os.rmdir(tmp)


if __name__ == "__main__":
    main()
