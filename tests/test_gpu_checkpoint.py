"""Checkpoints in the reference's binary formats and the rb3d host driver.

Restart must continue bit for bit (the reference recomputes F from phi on
restart exactly as the next step's forcing would, src/initial.cu:202).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from lbm_ddf_amd import Cavity

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
RB3D = ROOT / "lbm_ddf_amd" / "_lib" / "rb3d"
RB = dict(rayleigh=1.0e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)


def states_equal(a, b):
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("precision", ["fp64", "fp16", "bf16"])
def test_restart_continues_bit_for_bit(tmp_path, precision):
    shape = (14, 10, 9)
    a = Cavity(*shape, precision=precision, **RB)
    a.init_conductive(3, 1e-2)
    a.step(10)
    flow, fg = tmp_path / "000000010.dat", tmp_path / "fg000000010.dat"
    a.output_flow(flow)
    a.output_fg(fg)
    b = Cavity(*shape, precision=precision, **RB)
    b.restart(flow, fg, 10)
    assert b.step_count == 10
    a.step(7)
    b.step(7)
    states_equal(a, b)


def test_dump_layout_is_the_reference_layout(tmp_path):
    shape = (7, 6, 5)
    c = Cavity(*shape, **RB)
    c.init_conductive(1, 1e-2)
    c.step(3)
    c.output_flow(tmp_path / "flow.dat")
    c.output_fg(tmp_path / "fg.dat")
    n = 7 * 6 * 5
    flow = np.fromfile(tmp_path / "flow.dat", dtype="<f8")
    assert flow.size == 5 * n
    for i, k in enumerate(("rho", "ux", "uy", "uz", "phi")):
        assert np.array_equal(flow[i * n:(i + 1) * n], c.get_field(k).ravel()), k
    fg = np.fromfile(tmp_path / "fg.dat", dtype="<f8")
    assert fg.size == 2 * 27 * n
    assert np.array_equal(fg[:27 * n], c.get_pops("f").ravel())
    assert np.array_equal(fg[27 * n:], c.get_pops("g").ravel())


def run_rb3d(outdir, *extra):
    cmd = [str(RB3D), "--lx", "12", "--ly", "10", "--lz", "8", "--ra", "1e5", "--pr", "0.71", "--thot", "0.5",
           "--tcold", "-0.5", "--init", "conductive", "--seed", "4", "--amp", "0.01", "--ndiag", "20", "--nflowout",
           "20", "--nnuout", "20", "--outdir", str(outdir)] + list(extra)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_rb3d_driver_outputs_match_the_oracle(tmp_path, pattern):
    log = run_rb3d(tmp_path, "--nend", "40", "--pattern", pattern)
    assert "Simulation completed." in log and "Step 20" in log
    d = tmp_path / "Ra1.0e+05Pr0.71"
    o = oracle.OracleState(12, 10, 8, **RB)
    o.init_conductive(4, 0.01)
    nu_lines = (d / "Nu_1e50.txt").read_text().splitlines()
    assert len(nu_lines) == 2
    for line, step in zip(nu_lines, (20, 40)):
        o.run(step - (0 if step == 20 else 20))
        vals = [float(line[i:i + 15]) for i in range(0, 75, 15)]
        assert vals[0] == step and vals[1] == 0.0
        ref = o.output_nu()
        np.testing.assert_allclose(vals[2:], ref, rtol=1e-6, atol=1e-12)
    flow = np.fromfile(d / "000000040.dat", dtype="<f8").reshape(5, 8, 10, 12)
    for i, k in enumerate(("rho", "ux", "uy", "uz", "phi")):
        assert np.array_equal(flow[i], o[k]), k
    fg = np.fromfile(d / "fg000000040.dat", dtype="<f8").reshape(2, 27, 8, 10, 12)
    assert np.array_equal(fg[0], o["f"]) and np.array_equal(fg[1], o["g"])
    assert len((d / "statistics.dat").read_text().splitlines()) == 2
    prof = (d / "Profile_1e50.txt").read_text().splitlines()
    assert prof[0] == "20" and len(prof) == 2 * (1 + 10)


def test_rb3d_continue_equals_straight_run(tmp_path):
    """A checkpoint written by one storage pattern restarts in the other."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    run_rb3d(a, "--nend", "60")
    run_rb3d(b, "--nend", "40", "--pattern", "aa")
    run_rb3d(b, "--nend", "60", "--continue", "40")
    fa = (a / "Ra1.0e+05Pr0.71" / "fg000000060.dat").read_bytes()
    fb = (b / "Ra1.0e+05Pr0.71" / "fg000000060.dat").read_bytes()
    assert fa == fb
