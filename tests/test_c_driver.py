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
