"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

The oracle (oracle/oracle.c) restates src/ of the reference; the HIP FP64
kernels keep the reference's operation order and no FMA contraction, so the
stage kernels and the fused kernel are expected to agree with it bit for bit.
The acceptance bar of the FP64 path is 1e-12 relative max-norm on f, g, rho',
u and T after 1000 steps of the 64^3, Ra=1e5 cavity; the smaller cases below
check exact equality.
"""
import numpy as np
import pytest

import oracle
from lbm_ddf_amd import Cavity

pytestmark = pytest.mark.gpu

RB = dict(rayleigh=1.0e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
FIELDS = ("rho", "ux", "uy", "uz", "phi", "fy")


def make_pair(lx, ly, lz, seed=1, amp=1.0e-3, init="conductive", **kw):
    params = dict(RB)
    params.update(kw)
    c = Cavity(lx, ly, lz, **params)
    o = oracle.OracleState(lx, ly, lz, **params)
    if init == "conductive":
        c.init_conductive(seed, amp)
        o.init_conductive(seed, amp)
    else:
        c.init_arrays()
        o.init_arrays()
    return c, o


def gpu_state(c):
    s = c.state()
    s["fy"] = c.get_field("fy")
    return s


def rel_maxnorm(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def assert_state_equal(c, o, names=("f", "g") + FIELDS, exact=True, tol=1e-12):
    s = gpu_state(c)
    for k in names:
        a, b = s[k], o[k]
        if exact:
            bad = np.argwhere(a != b)
            assert bad.size == 0, f"{k}: {len(bad)} cells differ, first {bad[0]}: {a[tuple(bad[0])]!r} vs {b[tuple(bad[0])]!r}"
        else:
            r = rel_maxnorm(a, b)
            assert r <= tol, f"{k}: rel max-norm {r:.3e} > {tol}"


@pytest.mark.parametrize("init", ["conductive", "reference"])
@pytest.mark.parametrize("shape", [(5, 7, 9), (16, 12, 10)])
def test_initial_state_bitwise(shape, init):
    c, o = make_pair(*shape, init=init)
    assert_state_equal(c, o)


# A single cavity uses the row-blocked population layout for every shape;
# the plane layout (ensemble members) is compared with it in the ensemble tests.
@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 4, 3), (3, 3, 1), (70, 6, 5), (256, 4, 3)])
def test_each_stage_bitwise(shape):
    """Every stage entry point against the oracle stage, two full steps."""
    c, o = make_pair(*shape, seed=7, amp=1.0e-2)
    for _ in range(2):
        for stage in ("forcing", "collision_BGK", "collision_BGK_scalar", "streaming",
                      "streaming_scalar", "macrovar"):
            getattr(c, stage)()
            getattr(o, stage)()
            assert_state_equal(c, o)


@pytest.mark.parametrize("nsteps", [1, 2, 5, 23])
@pytest.mark.parametrize("shape", [(6, 5, 7), (65, 9, 4), (1, 4, 3), (3, 3, 1), (2, 5, 2), (130, 5, 3), (256, 5, 3)])
def test_fused_step_bitwise(shape, nsteps):
    c, o = make_pair(*shape, seed=3, amp=5.0e-2, rayleigh=1.0e6)
    c.step(nsteps)
    o.run(nsteps)
    assert c.step_count == nsteps
    assert_state_equal(c, o)


def test_fused_split_calls_and_reads_are_transparent():
    """step(3); read; step(4) is the same as step(7)."""
    c1, o = make_pair(12, 8, 6, seed=11, amp=1.0e-2)
    c2 = Cavity(12, 8, 6, **RB)
    c2.init_conductive(11, 1.0e-2)
    c1.step(3)
    c1.get_field("phi")
    c1.step(4)
    c2.step(7)
    o.run(7)
    assert_state_equal(c1, o)
    assert_state_equal(c2, o)


def test_graph_replays_between_reads_and_tuning_changes():
    """Long lbm_step calls replay captured runs of passes; the captured
    sequence depends on the buffer parity and launch shape, so reads (which
    complete the pending half-step), odd step counts and tuning changes
    between calls must leave the results exact."""
    c, o = make_pair(130, 9, 7, seed=13, amp=1.0e-2, rayleigh=1.0e6)
    done = 0
    for n, tuning in ((70, None), (97, (64, 3)), (64, None), (131, (256, 7)), (65, (128, 4))):
        if tuning:
            c.set_tuning(*tuning)
        c.step(n)
        done += n
        o.run(n)
        assert_state_equal(c, o)
    assert c.step_count == done


def test_fused_then_stages_then_fused():
    c, o = make_pair(9, 6, 5, seed=5, amp=1.0e-2)
    c.step(4)
    c.staged_step(2)
    c.step(3)
    o.run(9)
    assert_state_equal(c, o)


def test_config1_64cube_1000_steps():
    """64^3, Ra=1e5, Pr=0.71, walls +-0.5, seed 1, +-1e-3: 1000 steps.
    Tolerance: 1e-12 relative max-norm on f, g, rho', u, T."""
    c, o = make_pair(64, 64, 64, seed=1, amp=1.0e-3)
    c.step(1000)
    o.run(1000)
    assert_state_equal(c, o, names=("f", "g", "rho", "ux", "uy", "uz", "phi"), exact=False, tol=1e-12)
    # The FP64 path keeps the reference operation order: it is in fact exact.
    assert_state_equal(c, o, names=("f", "g", "rho", "ux", "uy", "uz", "phi"), exact=True)


def _diag_close(g, o, rtol=1e-11, atol=1e-13):
    for k, v in o.items():
        assert np.isclose(g[k], v, rtol=rtol, atol=atol), f"{k}: gpu {g[k]!r} oracle {v!r}"


def test_reduce_matches_oracle():
    c, o = make_pair(24, 20, 16, seed=2, amp=1.0e-2, rayleigh=1.0e6)
    c.step(60)
    o.run(60)
    _diag_close(c.reduce(), o.diag())


def test_run_diag_folded_matches_oracle():
    """Diagnostics folded into the fused launch, every 5 steps over 23 steps,
    against the oracle's output_nu / diag_flow at the same steps."""
    c, o = make_pair(20, 16, 12, seed=4, amp=1.0e-2, rayleigh=1.0e6)
    recs = c.run_diag(23, 5)
    assert [r["step"] for r in recs] == [5, 10, 15, 20]
    done = 0
    for r in recs:
        o.run(r["step"] - done)
        done = r["step"]
        _diag_close(r, o.diag())
    o.run(23 - done)
    assert_state_equal(c, o)
    # a second call continues from the finalized state
    recs2 = c.run_diag(7, 5)
    assert [r["step"] for r in recs2] == [25, 30]
    o.run(2)
    _diag_close(recs2[0], o.diag())


def test_run_diag_from_fused_phase():
    c, o = make_pair(10, 8, 6, seed=8, amp=1.0e-2)
    c.step(3)  # pending streaming half of step 3
    recs = c.run_diag(4, 2)
    assert [r["step"] for r in recs] == [4, 6]
    o.run(4)
    _diag_close(recs[0], o.diag())
    o.run(3)
    assert_state_equal(c, o)


def test_profile_matches_oracle():
    c, o = make_pair(14, 10, 12, seed=9, amp=1.0e-2, rayleigh=1.0e6)
    c.step(30)
    o.run(30)
    g, r = c.profile(), o.profile()
    # plane means: summation order differs (tree vs serial)
    np.testing.assert_allclose(g[:, :4], r[:, :4], rtol=1e-11, atol=1e-14)
    # T_rms = sqrt(<phi^2> - <phi>^2) cancels: compare against its scale
    np.testing.assert_allclose(g[:, 4], r[:, 4], rtol=0, atol=1e-12 * np.max(np.abs(r[:, 3])) / np.min(r[:, 4]))


@pytest.mark.parametrize("shape", [(7, 5, 6), (256, 3, 2)])
def test_pops_roundtrip_fp64(shape):
    c, _ = make_pair(*shape)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((27,) + shape[::-1])
    g = rng.standard_normal((27,) + shape[::-1])
    c.set_pops("f", f)
    c.set_pops("g", g)
    np.testing.assert_array_equal(c.get_pops("f"), f)
    np.testing.assert_array_equal(c.get_pops("g"), g)


@pytest.mark.parametrize("shape", [(32, 32, 32), (256, 24, 8)])
@pytest.mark.parametrize("prec,tol", [("fp16", 2e-3), ("bf16", 2e-2)])
def test_reduced_precision_tracks_fp64(prec, tol, shape):
    """16-bit storage with FP32 arithmetic follows the FP64 run (Ra=1e5,
    300 steps; both population layouts): max |dT| below tol, mass drift
    small, no NaN."""
    ref = Cavity(*shape, **RB)
    lp = Cavity(*shape, precision=prec, **RB)
    ref.init_conductive(1, 1.0e-3)
    lp.init_conductive(1, 1.0e-3)
    ref.step(300)
    lp.step(300)
    dT = np.max(np.abs(lp.get_field("phi") - ref.get_field("phi")))
    d = lp.reduce()
    assert d["nonfinite"] == 0
    assert dT < tol, f"max |dT| = {dT}"
    assert abs(d["mass"]) < 1e-1
    assert abs(d["nu_bulk"] - ref.reduce()["nu_bulk"]) < 0.05


def test_step_staged_reference_structure_bitwise():
    """lbm_step_staged (six launches + three syncs per step, src/main.cu:27-48)
    equals the oracle and the fused path."""
    c, o = make_pair(11, 7, 6, seed=21, amp=1.0e-2)
    c.step_staged(9)
    o.run(9)
    assert_state_equal(c, o)


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_non_default_physics_bitwise(pattern):
    """Parameters away from the defaults: gravity 0.5, an explicit beta, the
    reference's wall temperatures 1 and 0 (so T0 = 0.5 and the lagged force
    uses phi - 0.5), Pr = 7: fused steps and the reference step structure
    against the oracle, bit for bit."""
    params = dict(rayleigh=2.0e5, prandtl=7.0, t_hot=1.0, t_cold=0.0, gravity=0.5, beta=3.0e-3)
    c = Cavity(23, 9, 7, pattern=pattern, **params)
    o = oracle.OracleState(23, 9, 7, **params)
    c.init_conductive(9, 2.0e-2)
    o.init_conductive(9, 2.0e-2)
    c.step(13)
    o.run(13)
    assert_state_equal(c, o)
    c.step_staged(3)
    o.run(3)
    assert_state_equal(c, o)
    c.step(70)  # graph-replayed passes
    o.run(70)
    assert_state_equal(c, o)
