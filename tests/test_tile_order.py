"""The ensemble launch's tile order (lbm_ddf_amd/csrc/tile_order.hpp) on the
CPU: every tile is taken by exactly one block and each block's tile numbers
grow (so its loop ends), in the plain and the XCD-aware order, for one block
per tile, persistent grids and grids smaller than the XCD count."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("tile_order") / "tile_order_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", str(ROOT / "lbm_ddf_amd" / "csrc"),
                    "-o", str(exe), str(ROOT / "tests" / "tile_order_check.cpp")], check=True)
    return exe


def test_every_tile_taken_once(checker):
    r = subprocess.run([str(checker)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


def test_checker_catches_a_miscounted_xcd_share(checker):
    r = subprocess.run([str(checker), "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "FAIL" in r.stdout, r.stdout
