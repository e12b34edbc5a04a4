"""The oracle reproduces the committed golden fixtures bit for bit (CPU)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_fixture(name):
    meta = json.loads((GOLDEN / f"{name}.json").read_text())
    ref = np.load(GOLDEN / f"{name}.npz")
    o = oracle.OracleState(*meta["shape"], rayleigh=meta["rayleigh"], prandtl=meta["prandtl"],
                           t_hot=meta["t_hot"], t_cold=meta["t_cold"])
    if meta["init"] == "conductive":
        o.init_conductive(meta["seed"], meta["amp"])
    else:
        o.init_arrays()
    o.run(meta["steps"])
    for k in ref.files:
        assert np.array_equal(o[k], ref[k]), k
    for k, v in meta["diag"].items():
        assert o.diag()[k] == v, k


def test_fixtures_present():
    assert len(NAMES) >= 2
