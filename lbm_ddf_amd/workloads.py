"""Synthetic workloads of the measured configurations, shared by bench.py and
the GPU tests: the 64-cavity ensemble draw and the Poisson request stream of
the serving loop.  Draws come from numpy's PCG64 with fixed seeds, so every
caller gets the same cavities and requests."""
from __future__ import annotations

import math

import numpy as np


def ensemble_cavities(count: int = 64, draw_seed: int = 4, laterals=(32, 48, 64, 96, 128),
                      heights=(32, 64, 128), log_ra=(4.0, 7.0)) -> list:
    """Cavities of the ensemble workload: lateral side (x and z) from
    {32, 48, 64, 96, 128}, height (y) from {32, 64, 128}, Ra log-uniform in
    [1e4, 1e7], Pr from {0.71, 1.0, 4.4}; cavity k is seeded 1000 + k.  Other
    side, height and log10(Ra) ranges draw other ensembles with the same rules."""
    rng = np.random.default_rng(draw_seed)
    cavs = []
    for k in range(count):
        lat = int(rng.choice(list(laterals)))
        h = int(rng.choice(list(heights)))
        ra = float(10 ** rng.uniform(*log_ra))
        pr = float(rng.choice([0.71, 1.0, 4.4]))
        cavs.append(dict(lx=lat, ly=h, lz=lat, rayleigh=ra, prandtl=pr, seed=1000 + k))
    return cavs


# The small-cavity draw: many cavities whose single-cavity steps are bound by
# launch latency rather than HBM bandwidth (Ra up to 1e6, where tau stays
# clear of 1/2 at these heights).
SMALL_DRAW = dict(count=256, laterals=(16, 24, 32), heights=(16, 24, 32), log_ra=(4.0, 6.0))


def serve_requests(rate: float, arrivals_s: float, draw_seed: int = 5, max_steps: int = 50000,
                   nu_tol: float = 1.0e-3, window: int = 1000, sides=(32, 256), log_ra=(4.0, 8.0)) -> list:
    """Poisson arrivals at `rate` per second during `arrivals_s` seconds; cube
    side log-uniform in [32, 256], Ra log-uniform in [1e4, 1e8], Pr from
    {0.71, 1.0, 7.0}; a request stops when its running Nu moves by less than
    nu_tol over `window` steps, or at max_steps.  Other side and log10(Ra)
    ranges draw other request streams with the same rules."""
    rng = np.random.default_rng(draw_seed)
    reqs, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > arrivals_s:
            break
        side = int(round(math.exp(rng.uniform(math.log(sides[0]), math.log(sides[1])))))
        reqs.append(dict(lx=side, ly=side, lz=side, rayleigh=float(10 ** rng.uniform(*log_ra)),
                         prandtl=float(rng.choice([0.71, 1.0, 7.0])), seed=5000 + len(reqs), amp=1.0e-3,
                         arrival_s=t, max_steps=max_steps, nu_tol=nu_tol, window=window))
    return reqs
