"""The scheduling policy of lbm_serve on the CPU (no GPU).

lbm_ddf_amd/csrc/serve_loop.hpp holds the request loop as a template over a
backend and a clock; tests/serve_loop_cpu.cpp instantiates it with a
synthetic cavity backend and a virtual clock and checks admission (first
fit, HBM budget, resident caps, rejection), the convergence and budget
rules, divergence retirement, the time limit, arrivals and the statistics.
Built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer.
"""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_serve_loop_policy(tmp_path):
    exe = tmp_path / "serve_loop_cpu"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", f"-I{ROOT / 'include'}", f"-I{ROOT / 'lbm_ddf_amd' / 'csrc'}",
                    str(ROOT / "tests" / "serve_loop_cpu.cpp"), "-o", str(exe)], check=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
