"""Access-pattern copies (lbm_pattern_copy) against a numpy restatement of
each pass's addressing.

The copies carry no arithmetic, so every element must land, bit for bit,
where the rule puts it; unique element values make a wrong address visible.
The rules are the streaming of src/streaming.cu:3-251 in its gather form
(a population whose source lies outside the domain takes the cell's own
opposite population, the half-way bounce-back), its scatter form, and the
two in-place passes (the even pass writes population j into the slot of
opp(j); the odd pass reads (opp(i), x - c_i) and writes (j, x + c_j)).
Layout: element ((iz*ly + iy)*54 + s)*lx + ix, s = i for f_i and 27 + i for
g_i, i.e. a numpy array of shape (lz, ly, 54, lx).
"""
import numpy as np
import pytest
import torch

import oracle
from lbm_ddf_amd import capi

pytestmark = pytest.mark.gpu

CX, CY, CZ, _ = (np.asarray(t) for t in oracle.tables())
OPP = [int(np.flatnonzero((CX == -CX[i]) & (CY == -CY[i]) & (CZ == -CZ[i]))[0]) for i in range(27)]


def inside(x, y, z, shape):
    lz, ly, lx = shape
    return (x >= 0) & (x < lx) & (y >= 0) & (y < ly) & (z >= 0) & (z < lz)


def expected(pattern, a):
    lz, ly, _, lx = a.shape
    shape = (lz, ly, lx)
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    if pattern == "copy":
        return a.copy()
    if pattern == "inplace_even":
        out = a.copy()
        for s in range(54):
            i, base = s % 27, s - s % 27
            out[:, :, base + OPP[i], :] = a[:, :, s, :]
        return out
    out = np.zeros_like(a)
    written = np.zeros(a.shape, dtype=bool)
    for s in range(54):
        i, base = s % 27, s - s % 27
        o = base + OPP[i]
        # value of population s at each cell
        if pattern in ("pull", "inplace_odd") and i != 0:
            xs, ys, zs = x - CX[i], y - CY[i], z - CZ[i]
            ok = inside(xs, ys, zs, shape)
            frm = s if pattern == "pull" else o
            bounce = o if pattern == "pull" else s
            val = np.where(ok, a[np.clip(zs, 0, lz - 1), np.clip(ys, 0, ly - 1), frm, np.clip(xs, 0, lx - 1)],
                           a[z, y, bounce, x])
        else:
            val = a[:, :, s, :]
        # where it goes
        if pattern in ("push", "inplace_odd") and i != 0:
            xd, yd, zd = x + CX[i], y + CY[i], z + CZ[i]
            ok = inside(xd, yd, zd, shape)
            dz, dy, dx = np.where(ok, zd, z), np.where(ok, yd, y), np.where(ok, xd, x)
            ds = np.where(ok, s, o)
        else:
            dz, dy, dx, ds = z, y, x, np.full_like(x, s)
        assert not written[dz, dy, ds, dx].any()
        out[dz, dy, ds, dx] = val
        written[dz, dy, ds, dx] = True
    assert written.all()  # every slot is written exactly once
    return out


def unique_values(shape, elem_bytes, seed):
    n = 54 * int(np.prod(shape))
    if elem_bytes == 8:
        return (np.arange(n, dtype=np.float64) + 0.25 + seed).reshape(shape[0], shape[1], 54, shape[2])
    h = (np.arange(n, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(seed)) >> np.uint64(13)
    return (h & np.uint64(0xFFFF)).astype(np.uint16).view(np.int16).reshape(shape[0], shape[1], 54, shape[2])


SHAPES = [((4, 5, 70), 256, 7), ((3, 3, 512), 512, 7), ((3, 3, 512), 256, 3), ((3, 3, 1), 128, 7),
          ((4, 4, 130), 128, 3), ((5, 3, 65), 128, 7), ((3, 4, 300), 512, 3)]


@pytest.mark.parametrize("pattern", list(capi.ACCESS_PATTERNS))
@pytest.mark.parametrize("elem_bytes", [8, 2])
@pytest.mark.parametrize("shape,bx,flags", SHAPES)
def test_pattern_copy_moves_every_element_to_its_slot(pattern, elem_bytes, shape, bx, flags):
    lz, ly, lx = shape
    a = unique_values(shape, elem_bytes, seed=bx + flags)
    src = torch.from_numpy(a.copy()).to("cuda")
    inplace = pattern.startswith("inplace")
    dst = src if inplace else torch.full_like(src, -1)
    print(f"pattern {pattern} elem {elem_bytes} shape {lx}x{ly}x{lz} bx {bx} flags {flags}", flush=True)
    capi.pattern_copy(pattern, elem_bytes, (lx, ly, lz), src.data_ptr(), dst.data_ptr(), bx, flags)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    want = expected(pattern, a)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_pattern_bandwidth_times_a_pass():
    for pattern in capi.ACCESS_PATTERNS:
        ms = capi.pattern_bandwidth(pattern, 8, (64, 16, 16), 256, 7, reps=3)
        assert 0.0 < ms < 1000.0, pattern


@pytest.mark.parametrize("precision,pattern", [("fp64", "ab"), ("fp16", "ab"), ("fp64", "aa"), ("bf16", "aa")])
def test_pattern_time_on_a_cavity_leaves_it_uninitialised(precision, pattern):
    from lbm_ddf_amd import Cavity

    c = Cavity(70, 12, 8, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5, precision=precision, pattern=pattern)
    try:
        c.init_conductive(1, 1e-3)
        c.step(3)
        ms0, ms1 = c.pattern_time(3)
        assert 0.0 < ms0 < 1000.0 and 0.0 < ms1 < 1000.0
        with pytest.raises(RuntimeError, match="no state"):
            c.step(1)
        c.init_conductive(1, 1e-3)  # a fresh start works again ...
        c.step(4)
        got = c.reduce()
    finally:
        c.close()
    fresh = Cavity(70, 12, 8, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5, precision=precision,
                   pattern=pattern)
    try:
        fresh.init_conductive(1, 1e-3)
        fresh.step(4)
        want = fresh.reduce()
    finally:
        fresh.close()
    assert got == want  # ... and steps exactly like a new cavity


def test_pattern_time_refuses_an_ensemble_member():
    from lbm_ddf_amd import Ensemble

    e = Ensemble("fp64")
    try:
        m = e.add(24, 12, 8, seed=1, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
        with pytest.raises(RuntimeError, match="ensemble member"):
            e.cavity(m).pattern_time(2)
    finally:
        e.close()
