"""Row reference temperatures of the 16-bit g storage (CPU only).

A stored g value is decoded by a cell of the row below or above the one that
encoded it (pull scheme), or encoded for the row that will read it (in-place
pattern).  lbm_ddf_amd/csrc/tref.hpp gives every side the same binary32
T_ref(row); tests/tref_rows.cpp checks that for every row at heights that
are not powers of two, and that recomputing a neighbour's value in binary32
from the own row's (the rounding this replaces) would fail the same check;
and that writer_row() names, for every slot of the in-place pattern's
reference layout, the row of the cell whose write lands there (replaying
the odd pass's write rule on small grids).
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("tref") / "tref_rows"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{ROOT / 'lbm_ddf_amd' / 'csrc'}",
                    str(ROOT / "tests" / "tref_rows.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("walls", [("0.5", "-0.5"), ("1.0", "0.0"), ("0.7", "-0.3")])
def test_encoder_and_decoder_rows_agree(checker, walls):
    r = subprocess.run([str(checker), *walls, "96", "100", "127", "400", "512", "33"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "writer_rows mismatches 0", r.stdout
    fields = lines[1].split()
    pairs, bad, naive_bad = int(fields[1]), int(fields[3]), int(fields[5])
    assert pairs > 0 and bad == 0
    # the binary32 recomputation disagrees on a sizeable share of rows
    assert naive_bad > pairs // 20, r.stdout
