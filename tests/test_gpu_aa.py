"""The in-place (AA) pattern: one population buffer per distribution.

Even passes update each cell's own slots, odd passes its neighbours'; the
wall rules are applied by the writer.  The arithmetic is that of the A/B
kernel, so FP64 results equal the oracle (and the A/B path) bit for bit
whatever the parity of the step count at which the reference state is read.
"""
import numpy as np
import pytest

import oracle
from lbm_ddf_amd import Cavity, LbmError

from test_gpu_parity import RB, assert_state_equal, gpu_state, _diag_close

pytestmark = pytest.mark.gpu


def make_pair(lx, ly, lz, seed=1, amp=1.0e-3, init="conductive", **kw):
    params = dict(RB)
    params.update(kw)
    c = Cavity(lx, ly, lz, pattern="aa", **params)
    o = oracle.OracleState(lx, ly, lz, **params)
    if init == "conductive":
        c.init_conductive(seed, amp)
        o.init_conductive(seed, amp)
    else:
        c.init_arrays()
        o.init_arrays()
    return c, o


@pytest.mark.parametrize("nsteps", [1, 2, 3, 8, 23])
@pytest.mark.parametrize("shape", [(6, 5, 7), (65, 9, 4), (1, 4, 3), (3, 3, 1), (130, 5, 3), (256, 5, 3)])
def test_aa_fused_bitwise(shape, nsteps):
    c, o = make_pair(*shape, seed=3, amp=5.0e-2, rayleigh=1.0e6)
    c.step(nsteps)
    o.run(nsteps)
    assert c.step_count == nsteps
    assert_state_equal(c, o)


def test_aa_graph_replays_both_layouts():
    """Replayed runs of in-place passes starting from either layout."""
    c, o = make_pair(33, 7, 6, seed=17, amp=1.0e-2, rayleigh=1.0e6)
    for n in (65, 1, 96, 70):
        c.step(n)
        o.run(n)
        assert_state_equal(c, o)


def test_aa_split_calls_reads_and_stages():
    """Odd and even pass parities interleaved with reads and stage calls."""
    c, o = make_pair(12, 8, 6, seed=11, amp=1.0e-2)
    for k in (3, 4, 1, 2):
        c.step(k)
        o.run(k)
        assert_state_equal(c, o)
    c.staged_step(2)
    o.run(2)
    assert_state_equal(c, o)
    c.step(5)
    o.run(5)
    assert_state_equal(c, o)


@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 4, 3), (3, 3, 1), (70, 6, 5), (256, 4, 3)])
def test_aa_each_stage_bitwise(shape):
    """The stage interface in place (flip + exchange streaming) against the
    oracle's stages."""
    c, o = make_pair(*shape, seed=7, amp=1.0e-2)
    for _ in range(2):
        for stage in ("forcing", "collision_BGK", "collision_BGK_scalar", "streaming",
                      "streaming_scalar", "macrovar"):
            getattr(c, stage)()
            getattr(o, stage)()
            assert_state_equal(c, o)


def test_aa_step_staged_bitwise():
    c, o = make_pair(11, 7, 6, seed=21, amp=1.0e-2)
    c.step_staged(5)
    o.run(5)
    assert_state_equal(c, o)


def test_aa_equals_ab_64cube_1000_steps():
    """The 64^3, Ra=1e5 cavity: in place and A/B agree bit for bit after 1000
    steps (the A/B path equals the oracle there), and after 999."""
    ab = Cavity(64, 64, 64, **RB)
    aa = Cavity(64, 64, 64, pattern="aa", **RB)
    for c in (ab, aa):
        c.init_conductive(1, 1.0e-3)
        c.step(999)
    sa, sb = gpu_state(aa), gpu_state(ab)
    for k in sb:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    for c in (ab, aa):
        c.step(1)
    sa, sb = gpu_state(aa), gpu_state(ab)
    for k in sb:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_aa_run_diag_matches_ab_and_oracle():
    ab = Cavity(20, 16, 12, rayleigh=1.0e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    aa, o = make_pair(20, 16, 12, seed=4, amp=1.0e-2, rayleigh=1.0e6)
    ab.init_conductive(4, 1.0e-2)
    ra, rb = aa.run_diag(23, 5), ab.run_diag(23, 5)
    assert ra == rb
    o.run(20)
    _diag_close(ra[-1], o.diag())
    # diagnostics of a finalize from either layout
    for n in (1, 2):
        ra, rb = aa.run_diag(n, 1), ab.run_diag(n, 1)
        assert ra == rb


@pytest.mark.parametrize("bx,flags", [(64, 0), (64, 3), (128, 4), (256, 7), (256, 3)])
def test_aa_launch_variants_identical(bx, flags):
    ref, _ = make_pair(130, 6, 5, seed=2, amp=1.0e-2)
    c, _ = make_pair(130, 6, 5, seed=2, amp=1.0e-2)
    c.set_tuning(bx, flags)
    ref.step(7)
    c.step(7)
    sa, sb = gpu_state(c), gpu_state(ref)
    for k in sb:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_aa_rejects_unbuilt_flags_and_autotunes():
    c, o = make_pair(64, 8, 8, seed=5, amp=1.0e-2)
    with pytest.raises(LbmError):
        c.set_tuning(64, 1)
    t = c.autotune(1)
    assert t["flags"] in (0, 3, 4, 7)
    o.run(t["steps_taken"])
    assert_state_equal(c, o)


def test_aa_pops_roundtrip_and_restart(tmp_path):
    c, _ = make_pair(7, 5, 6)
    f = np.random.default_rng(0).standard_normal((27, 6, 5, 7))
    c.set_pops("f", f)
    np.testing.assert_array_equal(c.get_pops("f"), f)
    # checkpoint at an odd step count (layout E pending), restart, continue
    a, _ = make_pair(10, 8, 6, seed=6, amp=1.0e-2)
    straight, _ = make_pair(10, 8, 6, seed=6, amp=1.0e-2)
    a.step(7)
    a.output_flow(tmp_path / "flow.dat")
    a.output_fg(tmp_path / "fg.dat")
    b = Cavity(10, 8, 6, pattern="aa", **RB)
    b.restart(tmp_path / "flow.dat", tmp_path / "fg.dat", 7)
    b.step(6)
    straight.step(13)
    sa, sb = gpu_state(b), gpu_state(straight)
    for k in sb:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


@pytest.mark.parametrize("shape", [(32, 32, 32), (256, 24, 8)])
@pytest.mark.parametrize("prec,tol", [("fp16", 2e-3), ("bf16", 2e-2)])
def test_aa_reduced_precision_tracks_fp64(prec, tol, shape):
    """16-bit storage in place follows the FP64 run (Ra=1e5, 300 and 301
    steps; both population layouts) as closely as the A/B 16-bit path does:
    both encode g against the writer's row.  (Encoding against the reader's
    row, one row away, added w_i * dT / LY to every stored deviation and
    measured 2.3x A/B's max |dT| at LY = 24.)"""
    ref = Cavity(*shape, **RB)
    lp = Cavity(*shape, precision=prec, pattern="aa", **RB)
    lp_ab = Cavity(*shape, precision=prec, **RB)
    for c in (ref, lp, lp_ab):
        c.init_conductive(1, 1.0e-3)
    for n in (300, 1):
        for c in (ref, lp, lp_ab):
            c.step(n)
        phi = ref.get_field("phi")
        dT = np.max(np.abs(lp.get_field("phi") - phi))
        dT_ab = np.max(np.abs(lp_ab.get_field("phi") - phi))
        d = lp.reduce()
        assert d["nonfinite"] == 0
        assert dT < tol, f"max |dT| = {dT}"
        assert dT < 2.0 * dT_ab + 1e-6, (dT, dT_ab)
        assert abs(d["mass"]) < 1e-1
        assert abs(d["nu_bulk"] - ref.reduce()["nu_bulk"]) < 0.05
