"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

oracle/selftest.c drives every oracle entry point on small odd shapes
(LX = 1 and LZ = 1 included, where the wall blocks overlap most); built by
`make -C oracle sanitize` with -fsanitize=address,undefined, any invalid
access or undefined operation aborts the run.
"""
import pathlib
import shutil
import subprocess

import pytest

ORACLE = pathlib.Path(__file__).resolve().parent.parent / "oracle"


@pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None, reason="no C compiler")
def test_oracle_clean_under_sanitizers():
    subprocess.run(["make", "-C", str(ORACLE), "sanitize"], check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(ORACLE / "_build" / "selftest_sanitized")], capture_output=True, text=True, timeout=300,
                       env={"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if "->" in ln]
    assert len(lines) == 4 and all(ln.endswith("ok") for ln in lines), r.stdout
