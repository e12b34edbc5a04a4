"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

Each fixture is a tiny cavity run through the oracle (the C restatement of
the reference's src/ step): inputs are the parameters below, outputs the full
reference state (f, g, rho', u, phi) and the diagnostics after `steps` steps.
Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
import oracle  # noqa: E402

CASES = {
    "cavity_5x4x3_ra1e6_12steps": dict(shape=(5, 4, 3), rayleigh=1e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5,
                                       seed=7, amp=1e-2, init="conductive", steps=12),
    "cavity_4x6x5_reference_init_9steps": dict(shape=(4, 6, 5), rayleigh=1e7, prandtl=0.71, t_hot=1.0, t_cold=0.0,
                                               seed=0, amp=0.0, init="reference", steps=9),
}


def make(name, c):
    o = oracle.OracleState(*c["shape"], rayleigh=c["rayleigh"], prandtl=c["prandtl"], t_hot=c["t_hot"],
                           t_cold=c["t_cold"])
    if c["init"] == "conductive":
        o.init_conductive(c["seed"], c["amp"])
    else:
        o.init_arrays()
    o.run(c["steps"])
    np.savez_compressed(HERE / f"{name}.npz", **{k: o[k] for k in ("f", "g", "rho", "ux", "uy", "uz", "phi", "fy")})
    meta = dict(c)
    meta["diag"] = o.diag()
    (HERE / f"{name}.json").write_text(json.dumps(meta, indent=1))


if __name__ == "__main__":
    for name, case in CASES.items():
        make(name, case)
        print("wrote", name)
