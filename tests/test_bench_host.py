"""Host-side parts of bench.py (CPU only: nothing here starts a GPU run).

The bench line's `roofline.traffic` comes from profiles/traffic.txt, keyed by
workload and launch variant; the precision study's confidence intervals come
from batch means; a multi-GPU request must be refused with a message rather
than measured on one GPU and reported as N.
"""
import importlib.util
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_multi_gpu_request_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_table_has_the_default_workloads():
    b = load_bench()
    fp64 = b.traffic_from_profiles("fused_fp64_512")
    fp16 = b.traffic_from_profiles("fused_fp16_512")
    assert b.traffic_from_profiles("no_such_workload") is None
    # measured DRAM bytes per launch lie a few percent above the algorithmic bytes
    n = 512 ** 3
    assert 1.0 <= fp64 / (864 * n) < 1.10
    assert 1.0 <= fp16 / (216 * n) < 1.10
    # variant keys of the default launch shapes agree with the workload keys
    assert b.traffic_from_profiles("fused_fp64_512_bx256_f7") == fp64
    assert b.traffic_from_profiles("fused_fp16_512_bx256_f7") == fp16


@pytest.mark.parametrize("nbatch,t975", [(5, 2.776445), (10, 2.262157)])
def test_batch_means_interval(nbatch, t975):
    """The half-width is the Student-t interval of the batch means with
    nbatch - 1 degrees of freedom (tabulated t(0.975, 4) = 2.776445,
    t(0.975, 9) = 2.262157); a constant series has half-width 0."""
    b = load_bench()
    m, h = b.batch_means(np.full(100, 3.5), nbatch)
    assert m == 3.5 and h == 0.0
    rng = np.random.default_rng(0)
    x = rng.standard_normal(1000)
    m, h = b.batch_means(x, nbatch)
    means = np.array([v.mean() for v in np.array_split(x, nbatch)])
    assert m == pytest.approx(x.mean(), abs=1e-15)
    assert h == pytest.approx(t975 * means.std(ddof=1) / np.sqrt(nbatch), rel=1e-5)
    # independent normal samples: the interval covers the true mean 0 here
    assert abs(m) < h
