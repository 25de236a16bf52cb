"""The C ABI from C: include/axspmv.h compiles as C99 with -Wall -Werror, a
driver links libaxspmv.so and runs the host-side calls (no GPU needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpu_spmv_amd")


def build(tmp_path):
    exe = str(tmp_path / "driver")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_driver", "driver.c"), "-L", LIB, "-laxspmv",
                    f"-Wl,-rpath,{LIB}", "-lm", "-o", exe], check=True)
    return exe


def test_c_driver_host(tmp_path):
    out = subprocess.run([build(tmp_path)], check=True, capture_output=True, text=True).stdout
    assert "host: powerlaw n=20000" in out and "faxtuh1hw4" in out


@pytest.mark.gpu
def test_c_driver_gpu(tmp_path):
    out = subprocess.run([build(tmp_path), "gpu"], check=True, capture_output=True, text=True, timeout=120).stdout
    assert out.count("gpu: ") == 2, out


DRIVER = os.path.join(LIB, "axspmv_main")


def test_main_driver_arguments():
    """The C host driver (in place of the reference's main) rejects bad
    arguments, and without a HIP device it stops instead of running the
    kernels' rows on the CPU."""
    import gpu_spmv_amd as ax
    if not os.path.exists(DRIVER):
        pytest.skip("axspmv_main not built")
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode != 0 and "usage" in r.stderr
    r = subprocess.run([DRIVER, "laplace2d:8", "-q"], capture_output=True, text=True)
    assert r.returncode != 0 and "unknown argument" in r.stderr
    if ax.device_count() == 0:
        r = subprocess.run([DRIVER, "laplace2d:8"], capture_output=True, text=True)
        assert r.returncode != 0 and "no HIP device" in r.stderr


def parse_tables(out):
    """The two tables axspmv_main prints: formats (name -> mfp MB, beta,
    ct, device ct) and kernels (name -> exeTime, Gflops, ordTime, errAbs,
    errRel, rowInd, sErrRel)."""
    fm, km, sect = {}, {}, None
    for line in out.splitlines():
        if line.startswith("formats' data"):
            sect = "f"
        elif line.startswith("SpMV kernels' results"):
            sect = "k"
        elif sect and line.strip() and not line.startswith("---") and "memory [Mbytes]" not in line \
                and "exeTime" not in line:
            parts = line.split()
            if sect == "f":
                fm[" ".join(parts[:-4])] = parts[-4:]
            else:
                km[" ".join(parts[:-7])] = [float(v) for v in parts[-7:]]
    return fm, km


@pytest.mark.gpu
@pytest.mark.parametrize("spec", ["laplace2d:64", "random:3001:0:40", "file", "stencil27:12 -r"])
def test_main_driver_tables(tmp_path, spec):
    """axspmv_main on a generated matrix and on a .bin file: every format's
    row (mfp, beta, conversion times) and every kernel's row, each kernel
    within 1e-12 (signed) of the CPU CSR product; with -r also the
    renumbered RCT and AXT plans."""
    import gpu_spmv_amd as ax
    extra = 0
    if spec == "file":
        m = ax.powerlaw(5000, 2.0, 256, 12.0, seed=4)
        spec = str(tmp_path / "plaw.bin")
        m.write(spec)
        args = [DRIVER, spec, "-i", "20", "-x", "index", "-s"]
    elif spec.endswith(" -r"):
        args = [DRIVER, spec.split()[0], "-i", "20", "-r"]
        extra = 1
    else:
        args = [DRIVER, spec, "-i", "20", "-b", "128"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    fm, km = parse_tables(r.stdout)
    assert len(fm) == 16 + extra and len(km) == 21 + 2 * extra, (fm.keys(), km.keys())
    if extra:
        assert sum("_rcm]" in k for k in km) == 2, km.keys()
    assert any(k.startswith("faxtch1hw64bs") for k in fm)
    for name, row in km.items():
        et, gf, ot, ea, er, pos, ser = row
        assert et > 0 and gf > 0 and ser <= 1e-12, (name, row)
        assert (ot > 0) == ("ax layout" in name), (name, row)
    if spec.endswith(".bin"):
        assert os.path.exists(str(tmp_path / "plaw.sta"))
