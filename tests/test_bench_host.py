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


def test_cpu_baseline_runs_bound_oracle_in_a_child():
    """The CPU baseline times the oracle in a child process with OpenMP bound
    one thread per physical core, and reports the model and core count."""
    b = load_bench()
    r = b.cpu_baseline(12, 1e5, 3)
    assert r["kind"] == "port" and r["unit"] == "MLUPS" and r["value"] > 0
    assert r["cores"] >= 1 and "OMP_PROC_BIND=close" in r["sample"] and "12^3" in r["sample"]
    assert "cpu_model" in r and "physical_cores_available" in r


def test_precision_seeds_report_differences_and_intervals(tmp_path):
    """The seed report compares seed-averaged time means with FP64 and gives
    Welch 95% intervals from the seed-to-seed spread."""
    import json

    b = load_bench()
    rng = np.random.default_rng(0)

    def run(prec, shift):
        members = [{"seed": k + 1, "nu_bulk": 8.0 + shift + rng.normal(0, 0.2), "nu_top": 8.0 + shift,
                    "nu_bottom": 8.0 + shift, "kinetic_energy": 1e-3, "mass_drift_per_cell": 1e-6,
                    "nonfinite": 0.0, "samples": 100} for k in range(16)]
        return {"precision": prec, "size": 32, "ra": 1e6, "seeds": 16, "steps": 1000, "avg_from": 500,
                "sample": 100, "mlups": 1.0, "members": members}

    for prec, shift in (("fp64", 0.0), ("fp16", 0.04), ("bf16", 1.0)):
        (tmp_path / f"precision_seeds_{prec}_32_1e+06.json").write_text(json.dumps(run(prec, shift)))
    r = b.precision_seeds_report(32, 1e6, str(tmp_path))
    assert set(r) == {"fp64", "fp16", "bf16"}
    assert r["fp16"]["nu_top"]["diff_vs_fp64"] == pytest.approx(0.04)
    assert r["fp16"]["nu_top"]["rel_diff_vs_fp64"] == pytest.approx(0.005)
    assert r["bf16"]["nu_bulk"]["diff_vs_fp64"] - r["bf16"]["nu_bulk"]["diff_ci95"] > 0.5
    # 16 seeds with sd 0.2 each: a Welch half-width near 2.04 * 0.2 * sqrt(2/16)
    assert 0.08 < r["fp16"]["nu_bulk"]["diff_ci95"] < 0.25


def test_precision_seeds_report_pools_runs_with_seed_offsets(tmp_path):
    """Runs of one precision with different --seed-offset are pooled; a seed
    that appears twice is refused."""
    import json

    b = load_bench()

    def run(prec, first, n, shift):
        members = [{"seed": first + k, "nu_bulk": 8.0 + shift + 0.01 * k, "nu_top": 8.0 + shift,
                    "nu_bottom": 8.0 + shift, "kinetic_energy": 1e-3, "mass_drift_per_cell": 1e-6,
                    "nonfinite": 0.0, "samples": 10} for k in range(n)]
        return {"precision": prec, "size": 16, "ra": 1e5, "seeds": n, "first_seed": first, "steps": 100,
                "avg_from": 50, "sample": 10, "mlups": 2.0, "members": members}

    (tmp_path / "precision_seeds_fp64_16_100000.json").write_text(json.dumps(run("fp64", 1, 8, 0.0)))
    (tmp_path / "precision_seeds_fp16_16_100000.json").write_text(json.dumps(run("fp16", 1, 8, 0.1)))
    (tmp_path / "precision_seeds_fp16_16_100000_from9.json").write_text(json.dumps(run("fp16", 9, 8, 0.1)))
    r = b.precision_seeds_report(16, 1e5, str(tmp_path))
    assert r["fp16"]["seeds"] == 16 and r["fp16"]["runs"] == 2 and r["fp64"]["seeds"] == 8
    assert r["fp16"]["mlups"] == pytest.approx(2.0)
    assert r["fp16"]["nu_top"]["diff_vs_fp64"] == pytest.approx(0.1)
    (tmp_path / "precision_seeds_fp16_16_100000_from5.json").write_text(json.dumps(run("fp16", 5, 4, 0.1)))
    with pytest.raises(ValueError, match="same seed"):
        b.precision_seeds_report(16, 1e5, str(tmp_path))
