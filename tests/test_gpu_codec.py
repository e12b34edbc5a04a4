"""The 16-bit population codec, element by element (host copies through the C ABI).

FP16 storage keeps f and g - w_i*T_ref(iy) scaled by 2^15 as IEEE binary16;
BF16 storage keeps the same deviations as binary32 rounded to nearest even
to 16 bits.  T_ref(iy) = T_hot - dT*(iy + 1/2)/LY, rounded to binary32.  Host copies convert in
binary64 (a new value rounds through binary32 to the format), so numpy
reproduces every stored and decoded value exactly: set_pops followed by
get_pops must return the numpy emulation bit for bit, in both population
layouts, and writing back what was read changes nothing.
"""
import numpy as np
import pytest

import oracle
from lbm_ddf_amd import Cavity, Ensemble

pytestmark = pytest.mark.gpu

WALLS = dict(t_hot=0.5, t_cold=-0.5)
W = oracle.tables()[3]  # reference weights, src/initial.cu:77-83


def bf16_round(x32):
    """binary32 -> bfloat16 bits (round to nearest even, NaN kept quiet)."""
    u = x32.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return r.astype(np.uint32)


def bf16_to_f32(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def t_ref(shape):
    lz, ly, lx = shape
    dtref = (WALLS["t_hot"] - WALLS["t_cold"]) / ly
    tref0 = WALLS["t_hot"] - 0.5 * dtref
    # the binary32 value every kernel and the host codec use for a row
    tr = np.array([np.float32(tref0 - iy * dtref) for iy in range(ly)], dtype=np.float64)
    return tr[None, None, :, None]  # broadcast over (27, z, y, x)


def t_ref_writer(shape):
    """In-place storage: T_ref of the row of the cell that wrote each slot
    (i, z, y, x) of the reference state -- the source cell x - c_i when it
    lies inside the domain, else the cell itself (lbm_ddf_amd/csrc/tref.hpp)."""
    lz, ly, lx = shape
    cx, cy, cz, _ = oracle.tables()
    dtref = (WALLS["t_hot"] - WALLS["t_cold"]) / ly
    tref0 = WALLS["t_hot"] - 0.5 * dtref
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    out = np.empty((27,) + shape)
    for i in range(27):
        inside = ((x - cx[i] >= 0) & (x - cx[i] < lx) & (y - cy[i] >= 0) & (y - cy[i] < ly)
                  & (z - cz[i] >= 0) & (z - cz[i] < lz))
        row = np.where(inside, y - cy[i], y)
        out[i] = (tref0 - row * dtref).astype(np.float32).astype(np.float64)
    return out


def expected(prec, which, values, shape, writer_rows=False):
    """numpy emulation of encode-then-decode of the host copies."""
    w = W[:, None, None, None]
    tr = t_ref_writer(shape) if writer_rows else t_ref(shape)
    d = values - w * tr if which == "g" else values
    if prec == "fp16":
        h = (d * 32768.0).astype(np.float32).astype(np.float16)
        back = h.astype(np.float64) * (1.0 / 32768.0)
    else:
        back = bf16_to_f32(bf16_round(d.astype(np.float32))).astype(np.float64)
    return back + w * tr if which == "g" else back


def sample(shape, which, rng):
    """Deviations over several decades (the codec's working range), signed,
    on top of w_i*T for g."""
    z = (27,) + shape
    mag = 10.0 ** rng.uniform(-7, -0.3, size=z)
    dev = np.where(rng.random(z) < 0.5, -mag, mag)
    if which == "f":
        return dev
    tr = t_ref(shape).astype(np.float64)
    return W[:, None, None, None] * tr + dev


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("shape", [(20, 12, 10), (64, 6, 5)])
def test_codec_roundtrip_is_the_binary16_emulation(prec, shape):
    c = Cavity(*shape, precision=prec, **WALLS)
    c.init_conductive(1, 1.0e-3)
    rng = np.random.default_rng(3)
    zyx = shape[::-1]
    for which in ("f", "g"):
        vals = sample(zyx, which, rng)
        c.set_pops(which, vals)
        got = c.get_pops(which)
        exp = expected(prec, which, vals, zyx)
        bad = np.argwhere(got != exp)
        assert bad.size == 0, f"{prec} {which}: {len(bad)} values differ, first {bad[0]}: {got[tuple(bad[0])]!r} vs {exp[tuple(bad[0])]!r}"
        # the rounding is within half an ulp of the format (through binary32)
        dev = (vals - (W[:, None, None, None] * t_ref(zyx) if which == "g" else 0.0))
        rel = 2.0 ** -11 if prec == "fp16" else 2.0 ** -8
        assert np.all(np.abs(got - vals) <= (rel + 2.0 ** -24) * np.abs(dev) + 1e-15 * np.abs(vals) + 2.0 ** -40)
        # and writing back what was read changes no stored bit
        c.set_pops(which, got)
        np.testing.assert_array_equal(c.get_pops(which), got)
    c.close()


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_codec_roundtrip_plane_layout_ensemble_member(prec):
    """An ensemble member with a 20-cell x-line uses population planes."""
    e = Ensemble(prec)
    slot = e.add(20, 9, 7, seed=2, amp=1.0e-3, rayleigh=1e5, prandtl=0.71, **WALLS)
    m = e.cavity(slot)
    rng = np.random.default_rng(4)
    zyx = (7, 9, 20)
    for which in ("f", "g"):
        vals = sample(zyx, which, rng)
        m.set_pops(which, vals)
        np.testing.assert_array_equal(m.get_pops(which), expected(prec, which, vals, zyx))
    e.close()


@pytest.mark.parametrize("prec,rel", [("fp16", 2.0 ** -11), ("bf16", 2.0 ** -8)])
@pytest.mark.parametrize("nsteps,roundings", [(1, 2), (2, 4)])
def test_16bit_step_is_the_fp64_step_within_its_roundings(prec, rel, nsteps, roundings):
    """From the same stored state and fields, the 16-bit path (binary32
    arithmetic) and the FP64 oracle agree cell by cell to within the storage
    roundings on the way: one step stores each post-collision value (with
    its source row's reference temperature) and the finalize stores it again
    (with its own row's), two steps add the fused pass's pair.  Each rounding
    is at most half an ulp of the stored deviation, which differs between
    the two rows by w_i * dT / LY; binary32 arithmetic adds ~1e-7.  A wrong
    row temperature, weight or wall rule errs by w_i * dT / LY (>= 4e-4
    here), far above the bound."""
    shape = (20, 12, 10)
    zyx = shape[::-1]
    c = Cavity(*shape, precision=prec, rayleigh=1e6, prandtl=0.71, **WALLS)
    o = oracle.OracleState(*shape, rayleigh=1e6, prandtl=0.71, **WALLS)
    c.init_conductive(1, 1.0e-2)
    o.init_conductive(1, 1.0e-2)
    for k in ("rho", "ux", "uy", "uz", "phi", "fy"):
        np.testing.assert_array_equal(c.get_field(k), o[k])
    # the oracle starts from exactly the stored (rounded) populations
    o["f"] = c.get_pops("f")
    o["g"] = c.get_pops("g")
    c.step(nsteps)
    o.run(nsteps)
    dtref = (WALLS["t_hot"] - WALLS["t_cold"]) / shape[1]
    w = W[:, None, None, None]
    tr = t_ref(zyx).astype(np.float64)
    for which, dev, shift in (("f", o["f"], 0.0), ("g", o["g"] - w * tr, 2.0 * w * dtref)):
        got, exp = c.get_pops(which), o[which]
        bound = roundings * rel * (np.abs(dev) + shift) + 2e-7
        err = np.abs(got - exp)
        worst = np.unravel_index(np.argmax(err - bound), err.shape)
        assert np.all(err <= bound), f"{prec} {which}: error {err[worst]:.3e} > bound {bound[worst]:.3e} at {worst}"
    # phi is the FP32 sum of the decoded g: the same roundings, summed
    assert np.max(np.abs(c.get_field("phi") - o["phi"])) < 27 * roundings * rel * 0.02 + 1e-6


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_codec_roundtrip_in_place_writer_rows(prec):
    """In place, the reference state encodes g of each slot against the row of
    the cell that streamed it in (its own row for a wall value): set_pops then
    get_pops equals the numpy emulation of that rule bit for bit, and writing
    back what was read changes nothing."""
    shape = (20, 12, 10)
    zyx = shape[::-1]
    c = Cavity(*shape, precision=prec, pattern="aa", **WALLS)
    c.init_conductive(1, 1.0e-3)
    rng = np.random.default_rng(5)
    vals = sample(zyx, "g", rng)
    c.set_pops("g", vals)
    got = c.get_pops("g")
    np.testing.assert_array_equal(got, expected(prec, "g", vals, zyx, writer_rows=True))
    c.set_pops("g", got)
    np.testing.assert_array_equal(c.get_pops("g"), got)
    c.close()
