"""Ensemble (many cavities, one launch per step) against standalone cavities.

Each cavity's arithmetic in the ensemble kernel is the single-cavity
step_cell(), so after k of its own steps a member must equal a standalone
run of k steps bit for bit -- alone and in the mix, across hipGraph replays
with eager work between them, and across joins and removals.
"""
import numpy as np
import pytest

from lbm_ddf_amd import Cavity, Ensemble, LbmError

pytestmark = pytest.mark.gpu

CAVS = [
    dict(shape=(12, 10, 8), rayleigh=1e5, prandtl=0.71, seed=11),
    dict(shape=(16, 9, 11), rayleigh=3e5, prandtl=1.0, seed=12),
    dict(shape=(7, 12, 5), rayleigh=1e6, prandtl=4.4, seed=13),
    dict(shape=(20, 16, 14), rayleigh=5e4, prandtl=0.71, seed=14),
    dict(shape=(9, 8, 9), rayleigh=2e6, prandtl=1.0, seed=15),
    dict(shape=(256, 5, 3), rayleigh=1e6, prandtl=0.71, seed=16),
]
# Members whose x-line is a whole number of 128-byte lines (256 cells; 16 in
# FP64) are row-blocked, the others use population planes; the standalone
# cavities they are compared with are always row-blocked.

WALLS = dict(t_hot=0.5, t_cold=-0.5)
AMP = 1.0e-2


def standalone(cav, nsteps, precision="fp64", pattern="ab"):
    c = Cavity(*cav["shape"], rayleigh=cav["rayleigh"], prandtl=cav["prandtl"], precision=precision, pattern=pattern,
               **WALLS)
    c.init_conductive(cav["seed"], AMP)
    c.step(nsteps)
    return c


def build(precision="fp64", cavs=CAVS, pattern="ab"):
    e = Ensemble(precision)
    slots = [e.add(*c["shape"], seed=c["seed"], amp=AMP, rayleigh=c["rayleigh"], prandtl=c["prandtl"],
                   pattern=pattern, **WALLS)
             for c in cavs]
    return e, slots


def assert_member_equals(e, slot, cav, precision="fp64", pattern="ab"):
    m = e.cavity(slot)
    k = m.step_count
    ref = standalone(cav, k, precision, pattern)
    a, b = m.state(), ref.state()
    for key in a:
        assert np.array_equal(a[key], b[key]), f"slot {slot} {key} differs after {k} steps"
    return k


@pytest.mark.parametrize("graph_pairs,cpt,bpc", [(0, 1, 4), (4, 1, 4), (4, 2, 4), (0, 4, 4), (4, 1, 0), (0, 2, 0)])
def test_members_equal_standalone_runs(graph_pairs, cpt, bpc):
    """bpc: blocks per CU of the persistent grid, 0 for one block per tile."""
    e, slots = build()
    e.set_tuning(128, bpc, graph_pairs, 7 if graph_pairs else 0, cpt)
    e.step(37)
    for s, cav in zip(slots, CAVS):
        # prologue (collides step 1) + 37 launches
        assert assert_member_equals(e, s, cav) == 38
    e.close()


def test_graph_replays_with_eager_work_between():
    """Replays of the captured graph interleaved with host reads (which
    finalize members) and direct launches: every replay stays exact."""
    e, slots = build()
    e.set_tuning(64, 2, 3, 4)  # 6 steps per replay
    total = 0
    for n in (12, 7, 18):
        e.step(n)
        total += n + 1  # members were read (reference state): prologue + n launches
        ks = [assert_member_equals(e, s, c) for s, c in zip(slots, CAVS)]
        assert ks == [total] * len(slots)
    e.close()


@pytest.mark.parametrize("cpt", [1, 4])
def test_run_diag_matches_standalone_reduce(cpt):
    e, slots = build()
    e.set_tuning(64, 8, 4, 7, cpt)
    e.step(10)
    diags = e.run_diag(5)
    for s, cav in zip(slots, CAVS):
        d = diags[s]
        ref = standalone(cav, d["step"])
        r = ref.reduce()
        for k in ("nu_bulk", "nu_top", "nu_bottom", "umean", "vmean", "wmean", "tmean", "kinetic_energy",
                  "mass", "max_speed"):
            assert np.isclose(d[k], r[k], rtol=1e-11, atol=1e-14), (s, k, d[k], r[k])
        assert d["nonfinite"] == 0
    e.close()


def test_join_and_leave_between_calls():
    e, slots = build(cavs=CAVS[:3])
    e.step(9)
    s4 = e.add(*CAVS[3]["shape"], seed=CAVS[3]["seed"], amp=AMP, rayleigh=CAVS[3]["rayleigh"],
               prandtl=CAVS[3]["prandtl"], **WALLS)
    e.step(8)
    e.remove(slots[1])
    s5 = e.add(*CAVS[4]["shape"], seed=CAVS[4]["seed"], amp=AMP, rayleigh=CAVS[4]["rayleigh"],
               prandtl=CAVS[4]["prandtl"], **WALLS)
    assert s5 == slots[1]  # the freed slot is reused
    e.step(11)
    assert e.count()["resident"] == 4
    assert assert_member_equals(e, slots[0], CAVS[0]) == 29  # prologue + 9 + 8 + 11
    assert assert_member_equals(e, slots[2], CAVS[2]) == 29
    assert assert_member_equals(e, s4, CAVS[3]) == 20  # joined: prologue + 8 + 11
    assert assert_member_equals(e, s5, CAVS[4]) == 12  # prologue + 11
    e.close()


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_reduced_precision_members_equal_standalone(precision):
    e, slots = build(precision)
    e.step(21)
    for s, cav in zip(slots, CAVS):
        assert_member_equals(e, s, cav, precision)
    e.close()


def test_autotune_and_cache(tmp_path):
    e, slots = build()
    cache = str(tmp_path / "tune.txt")
    ms = e.autotune(4, cache)
    assert ms > 0
    t1 = e.tuning
    lines = open(cache).read().split("\n")
    assert len([x for x in lines if x.strip()]) == 1
    e2, _ = build()
    ms2 = e2.autotune(4, cache)  # served from the cache: same choice, no new line
    assert e2.tuning == t1
    assert ms2 == pytest.approx(ms, rel=1e-12)
    assert len([x for x in open(cache).read().split("\n") if x.strip()]) == 1
    # results are unaffected by the tuning
    for s, cav in zip(slots, CAVS):
        assert_member_equals(e, s, cav)
    e.close()
    e2.close()


def test_injected_nan_stays_in_its_cavity():
    """Host-forced fault: a NaN written into one member's populations shows up
    in that member's health record only; the other members still equal
    their standalone runs bit for bit."""
    e, slots = build()
    e.step(5)
    bad = e.cavity(slots[1])
    f = bad.get_pops("f")
    f[13, 1, 2, 3] = np.nan
    bad.set_pops("f", f)
    diags = e.run_diag(6)
    assert diags[slots[1]]["nonfinite"] > 0
    for s, cav in zip(slots, CAVS):
        if s == slots[1]:
            continue
        assert diags[s]["nonfinite"] == 0
        assert_member_equals(e, s, cav)
    e.close()


@pytest.mark.parametrize("graph_pairs,cpt", [(0, 1), (3, 1), (3, 4)])
def test_in_place_members_equal_standalone_runs(graph_pairs, cpt):
    """In-place members (one buffer, even and odd passes per cavity, from its
    own layout) equal standalone runs of either pattern bit for bit, across
    odd step counts, graph replays and host reads between calls."""
    e, slots = build(pattern="aa")
    e.set_tuning(128, 4, graph_pairs, 7 if graph_pairs else 0, cpt)
    total = 0
    for n in (13, 14):
        e.step(n)
        total += n + 1  # reads below finalize the members: a prologue per call
        for s, cav in zip(slots, CAVS):
            assert assert_member_equals(e, s, cav, pattern="aa") == total
            assert assert_member_equals(e, s, cav, pattern="ab") == total
    e.close()


def test_in_place_members_join_leave_and_diag():
    """Members at different step parities share the launches: a cavity that
    joins later runs its even pass while the others run their odd one."""
    e, slots = build(cavs=CAVS[:3], pattern="aa")
    e.set_tuning(64, 4, 2, 7)
    e.step(9)
    s4 = e.add(*CAVS[3]["shape"], seed=CAVS[3]["seed"], amp=AMP, rayleigh=CAVS[3]["rayleigh"],
               prandtl=CAVS[3]["prandtl"], pattern="aa", **WALLS)
    e.step(8)
    e.remove(slots[1])
    diags = e.run_diag(5)
    assert assert_member_equals(e, slots[0], CAVS[0], pattern="aa") == 23  # prologue + 9 + 8 + 5
    assert assert_member_equals(e, s4, CAVS[3], pattern="aa") == 14  # prologue + 8 + 5
    for s, cav in ((slots[0], CAVS[0]), (s4, CAVS[3])):
        d = diags[s]
        r = standalone(cav, d["step"], pattern="aa").reduce()
        for k in ("nu_bulk", "nu_top", "nu_bottom", "kinetic_energy", "mass"):
            assert np.isclose(d[k], r[k], rtol=1e-11, atol=1e-14), (s, k, d[k], r[k])
    e.close()


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_in_place_reduced_precision_members(precision):
    """16-bit in place: equal to the standalone in-place run bit for bit (the
    A/B run rounds g against other rows' reference temperatures, so it is
    close, not equal)."""
    e, slots = build(precision, pattern="aa")
    e.step(21)
    for s, cav in zip(slots, CAVS):
        assert_member_equals(e, s, cav, precision, pattern="aa")
    e.close()


def test_members_share_one_storage_pattern():
    e, slots = build(cavs=CAVS[:2])
    c = CAVS[2]
    with pytest.raises(RuntimeError, match="pattern"):
        e.add(*c["shape"], seed=c["seed"], amp=AMP, rayleigh=c["rayleigh"], prandtl=c["prandtl"], pattern="aa",
              **WALLS)
    for s in slots:
        e.remove(s)
    s = e.add(*c["shape"], seed=c["seed"], amp=AMP, rayleigh=c["rayleigh"], prandtl=c["prandtl"], pattern="aa",
              **WALLS)  # an empty ensemble takes the new pattern
    e.step(6)
    assert assert_member_equals(e, s, c, pattern="aa") == 7
    e.close()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_members_equal_oracle(pattern):
    """Every member of the odd-shaped set (population planes and row-blocked
    lines, x-lines of 7 to 256 cells) against the FP64 oracle directly."""
    import oracle

    e, slots = build(pattern=pattern)
    e.set_tuning(64, 4, 2, 7, 2)
    e.step(17)
    for s, cav in zip(slots, CAVS):
        m = e.cavity(s)
        assert m.step_count == 18  # prologue + 17 launches
        o = oracle.OracleState(*cav["shape"], rayleigh=cav["rayleigh"], prandtl=cav["prandtl"], **WALLS)
        o.init_conductive(cav["seed"], AMP)
        o.run(18)
        st = m.state()
        for key in ("f", "g", "rho", "ux", "uy", "uz", "phi"):
            assert np.array_equal(st[key], o[key]), (cav["shape"], pattern, key)
    e.close()


def test_member_cannot_be_stepped_or_destroyed_alone():
    """A member's buffers and parity belong to the ensemble's descriptor
    table: the calls that advance or free a context refuse a member, and the
    ensemble stays exact afterwards."""

    e, slots = build(cavs=CAVS[:2])
    e.step(4)
    m = e.cavity(slots[0])
    for call in (lambda: m.step(1), lambda: m.step_timed(1), lambda: m.run_diag(2, 1), lambda: m.autotune(1),
                 lambda: m.step_staged(1), lambda: m.forcing(), lambda: m.streaming()):
        with pytest.raises(LbmError, match="ensemble member"):
            call()
    assert m.lib.lbm_destroy(m.h) == -4  # LBM_E_STATE
    e.step(3)
    for s, cav in zip(slots, CAVS[:2]):
        assert assert_member_equals(e, s, cav) == 8  # prologue + 4 + 3: the refused calls read nothing
    e.close()


def test_one_block_per_tile_across_joins_and_graph_replays():
    """With one block per tile the launch grid follows the tile count: a
    join or a removal between graph replays re-captures the graph, and every
    member stays exact."""
    e, slots = build(cavs=CAVS[:3])
    e.set_tuning(64, 0, 2, 7, 1)  # 4 steps per replay
    e.step(9)
    s4 = e.add(*CAVS[5]["shape"], seed=CAVS[5]["seed"], amp=AMP, rayleigh=CAVS[5]["rayleigh"],
               prandtl=CAVS[5]["prandtl"], **WALLS)
    e.step(8)
    e.remove(slots[1])
    e.step(12)
    assert assert_member_equals(e, slots[0], CAVS[0]) == 30  # prologue + 9 + 8 + 12
    assert assert_member_equals(e, slots[2], CAVS[2]) == 30
    assert assert_member_equals(e, s4, CAVS[5]) == 21  # prologue + 8 + 12
    e.close()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_16bit_addressing_follows_the_largest_member(pattern):
    """16-bit members (A/B or in place) are addressed by 32-bit byte offsets while every
    member's buffer spans under 4 GiB, and by 64-bit ones otherwise: a member
    beyond that size joining and leaving switches the launch (and its
    captured graph) between the two, and every member stays exact.  (The
    small members are read only at the end: reading a 16-bit member stores
    its reference state, a rounding that a continuous run does not make.)"""
    big = dict(shape=(352, 352, 328), rayleigh=1e5, prandtl=0.71, seed=21)  # 54 x 40.6 M cells x 2 B > 4 GiB
    print("16-bit ensemble with a member of", big["shape"], "pattern", pattern, flush=True)
    e, slots = build("fp16", cavs=CAVS[:2], pattern=pattern)
    e.set_tuning(128, 0, 2, 7, 1)  # 4 steps per graph replay
    e.step(6)
    sb = e.add(*big["shape"], seed=big["seed"], amp=AMP, rayleigh=big["rayleigh"], prandtl=big["prandtl"],
               pattern=pattern, **WALLS)
    e.step(5)
    m = e.cavity(sb)
    k = m.step_count
    assert k == 6  # prologue + 5
    ref = standalone(big, k, "fp16", pattern)
    for name in ("phi", "uy"):  # the populations would take 17 GB of host memory in FP64
        assert np.array_equal(m.get_field(name), ref.get_field(name)), name
    ref.close()
    e.remove(sb)
    e.step(7)
    for s, cav in zip(slots, CAVS[:2]):
        assert assert_member_equals(e, s, cav, "fp16", pattern) == 19  # prologue + 6 + 5 + 7
    e.close()


def test_member_beyond_two_to_the_31_cells_is_refused_before_allocating():
    e = Ensemble("fp16")
    with pytest.raises(LbmError, match="2\\^31"):
        e.add(1291, 1291, 1289, seed=1, amp=AMP, rayleigh=1e5, prandtl=0.71, pattern="aa", **WALLS)
    s = e.add(*CAVS[0]["shape"], seed=CAVS[0]["seed"], amp=AMP, rayleigh=CAVS[0]["rayleigh"],
              prandtl=CAVS[0]["prandtl"], pattern="aa", **WALLS)
    e.step(3)
    assert assert_member_equals(e, s, CAVS[0], "fp16", "aa") == 4
    e.close()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_16bit_member_just_under_4gib_is_exact(pattern):
    """A 16-bit member whose buffer spans just under 4 GiB (54 x 39,766,464
    planes x 2 B = 4,294,778,112 B) is addressed by 32-bit byte offsets up to
    within 189 KB of 2^32; it must equal its standalone run."""
    near = dict(shape=(340, 340, 344), rayleigh=1e5, prandtl=0.71, seed=22)
    print("16-bit member of", near["shape"], "pattern", pattern, flush=True)
    e = Ensemble("fp16")
    s = e.add(*near["shape"], seed=near["seed"], amp=AMP, rayleigh=near["rayleigh"], prandtl=near["prandtl"],
              pattern=pattern, **WALLS)
    e.set_tuning(256, 0, 2, 7, 1)
    e.step(9)
    m = e.cavity(s)
    assert m.step_count == 10  # prologue + 9
    ref = standalone(near, 10, "fp16", pattern)
    for name in ("phi", "uy", "rho"):
        assert np.array_equal(m.get_field(name), ref.get_field(name)), name
    ref.close()
    e.close()
