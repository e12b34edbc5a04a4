"""The ensemble kernel's multiply-shift division of cell indices
(lbm_ddf_amd/csrc/div_magic.hpp) on the CPU, against the integer division,
with a planted mistake that the check must catch."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("div_magic") / "div_magic_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", str(ROOT / "lbm_ddf_amd" / "csrc"),
                    "-o", str(exe), str(ROOT / "tests" / "div_magic_check.cpp")], check=True)
    return exe


def test_division_equals_integer_division(checker):
    r = subprocess.run([str(checker)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


def test_checker_catches_a_rounded_down_multiplier(checker):
    r = subprocess.run([str(checker), "1"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "FAIL" in r.stdout, r.stdout
