"""The HIP path (fused and staged) reproduces the golden fixtures bit for bit,
without using the oracle at run time."""
import json
from pathlib import Path

import numpy as np
import pytest

from lbm_ddf_amd import Cavity

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


@pytest.mark.parametrize("path", ["fused", "staged"])
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(name, path):
    meta = json.loads((GOLDEN / f"{name}.json").read_text())
    ref = np.load(GOLDEN / f"{name}.npz")
    c = Cavity(*meta["shape"], rayleigh=meta["rayleigh"], prandtl=meta["prandtl"], t_hot=meta["t_hot"],
               t_cold=meta["t_cold"])
    if meta["init"] == "conductive":
        c.init_conductive(meta["seed"], meta["amp"])
    else:
        c.init_arrays()
    if path == "fused":
        c.step(meta["steps"])
    else:
        c.staged_step(meta["steps"])
    s = c.state()
    s["fy"] = c.get_field("fy")
    for k in ref.files:
        assert np.array_equal(s[k], ref[k]), k
    d = c.reduce()
    for k, v in meta["diag"].items():
        assert np.isclose(d[k], v, rtol=1e-12, atol=1e-15), (k, d[k], v)
