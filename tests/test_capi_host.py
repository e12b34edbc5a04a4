"""Host-side checks of the C ABI (CPU only: no call touches a GPU).

The library must load, export every function declared in include/lbm_ddf.h,
and its host-only entry points (parameter derivation, sizing, argument
validation) must agree with the oracle and the reference's formulas.
"""
import ctypes
import re
from pathlib import Path

import pytest

import oracle
from lbm_ddf_amd import capi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "lbm_ddf.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-zA-Z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes binding covers them all
    assert set(names) == set(capi.SIGNATURES)


@pytest.mark.parametrize("L,ra,pr", [(64, 1e5, 0.71), (512, 1e7, 0.71), (512, 1e8, 0.71), (37, 3e4, 4.4)])
def test_derive_matches_oracle(L, ra, pr):
    d = capi.derive(capi.make_params(L, L, L, rayleigh=ra, prandtl=pr, t_hot=0.5, t_cold=-0.5))
    o = oracle.make_params(L, L, L, rayleigh=ra, prandtl=pr, t_hot=0.5, t_cold=-0.5)
    for k in ("visc", "tau", "u0", "diff", "tauc", "t0"):
        assert d[k] == getattr(o, k), k


def test_reference_default_parameters():
    """src/parameters.h: 400^3, Ra=1e7, Pr=0.71, T_HOT=1, T_COLD=0, g=1."""
    d = capi.derive(capi.make_params(400, 400, 400))
    assert d["t0"] == 0.5
    assert d["tau"] == pytest.approx(3 * 400 * (0.02 * 0.71 / 1e7) ** 0.5 + 0.5, rel=1e-14)


def test_bytes_required_fp64_512():
    b = capi.bytes_required(capi.make_params(512, 512, 512))
    n = 512 ** 3
    assert b >= 4 * 27 * n * 8 + 6 * n * 8
    assert b < 125e9  # fits one 288 GB MI355X with room to spare
    b16 = capi.bytes_required(capi.make_params(512, 512, 512, precision="fp16"))
    assert b16 < b / 2.5


def test_bytes_required_in_place_pattern():
    """One population buffer per distribution instead of two: the 512^3 FP64
    cavity needs about 65 GB instead of 124 GB."""
    ab = capi.bytes_required(capi.make_params(512, 512, 512))
    aa = capi.bytes_required(capi.make_params(512, 512, 512, pattern="aa"))
    n = 512 ** 3
    assert ab - aa == 54 * n * 8 + 2 * 512 * 512 * 8  # the second f+g buffer and the wall-rho array
    assert aa < 66e9


@pytest.mark.parametrize("shape", [(4, 2, 4), (0, 8, 8), (8, 8, 0)])
def test_create_rejects_bad_geometry_without_touching_the_gpu(shape):
    lib = capi.load()
    h = ctypes.c_void_p()
    p = capi.make_params(*shape)
    rc = lib.lbm_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1
    assert b"ly >= 3" in lib.lbm_last_error()
    assert not h.value


def test_create_rejects_bad_precision_and_physics():
    lib = capi.load()
    h = ctypes.c_void_p()
    p = capi.make_params(8, 8, 8)
    p.precision = 7
    assert lib.lbm_create(ctypes.byref(p), ctypes.byref(h)) == -1
    p = capi.make_params(8, 8, 8, rayleigh=-1.0)
    assert lib.lbm_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"Ra" in lib.lbm_last_error()
    p = capi.make_params(8, 8, 8)
    p.pattern = 2
    assert lib.lbm_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"pattern" in lib.lbm_last_error()
    p = capi.make_params(8, 8, 8)
    p.reserved = 1
    assert lib.lbm_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"reserved" in lib.lbm_last_error()
    assert not h.value


@pytest.mark.parametrize("walls", [(1.5, 0.0), (0.0, -2.0), (float("nan"), 0.0)])
def test_fp16_storage_rejects_wall_temperatures_beyond_its_range(walls):
    """Scaled binary16 holds magnitudes below 2: FP16 storage takes wall
    temperatures in [-1, 1] (the reference default is 1 and 0); FP64 and BF16
    take any."""
    lib = capi.load()
    h = ctypes.c_void_p()
    p = capi.make_params(8, 8, 8, precision="fp16", t_hot=walls[0], t_cold=walls[1])
    assert lib.lbm_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"FP16" in lib.lbm_last_error()
    assert not h.value
    nbytes = ctypes.c_uint64()
    for prec in ("fp64", "bf16"):
        q = capi.make_params(8, 8, 8, precision=prec, t_hot=walls[0], t_cold=walls[1])
        assert lib.lbm_bytes_required(ctypes.byref(q), ctypes.byref(nbytes)) == 0
    q = capi.make_params(8, 8, 8, precision="fp16", t_hot=1.0, t_cold=-1.0)
    assert lib.lbm_bytes_required(ctypes.byref(q), ctypes.byref(nbytes)) == 0


def test_null_context_is_an_error_not_a_crash():
    lib = capi.load()
    assert lib.lbm_step(None, 1) == -1
    assert lib.lbm_finalize(None) == -1
    assert lib.lbm_destroy(None) == 0


def test_argument_errors_of_the_other_entry_points_need_no_gpu():
    """Validation happens before any device work: bad arguments come back as
    LBM_E_INVALID with a message, never as a crash."""
    lib = capi.load()
    v = ctypes.c_double()
    assert lib.lbm_copy_bandwidth(ctypes.c_uint64(100), 5, ctypes.byref(v)) == -1
    assert lib.lbm_copy_bandwidth(ctypes.c_uint64(1 << 20), 0, ctypes.byref(v)) == -1
    e = ctypes.c_void_p()
    assert lib.lbm_ensemble_create(9, ctypes.byref(e)) == -1
    assert b"precision" in lib.lbm_last_error()
    assert not e.value
    assert lib.lbm_ensemble_destroy(None) == 0
    # the request loop checks its options first
    reqs = (capi.Request * 1)()
    reqs[0].lx = reqs[0].ly = reqs[0].lz = 8
    reqs[0].window = 100
    reqs[0].max_steps = 100
    res = (capi.Result * 1)()
    st = capi.ServeStats()

    def run(**kw):
        o = dict(precision=0, max_resident=0, chunk_steps=50, hbm_budget=1 << 30, time_limit_s=1.0, block_x=0,
                 blocks_per_cu=0, flags=7, cells_per_thread=1, max_resident_cells=0, progress_s=0.0, pattern=0,
                 reserved=0)
        o.update(kw)
        opt = capi.ServeOptions(**o)
        return lib.lbm_serve(reqs, 1, ctypes.byref(opt), res, ctypes.byref(st))

    assert run(chunk_steps=0) == -1 and b"chunk_steps" in lib.lbm_last_error()
    assert run(chunk_steps=30) == -1 and b"multiple of chunk_steps" in lib.lbm_last_error()
    assert run(pattern=2) == -1 and b"pattern" in lib.lbm_last_error()
    assert run(reserved=1) == -1 and b"reserved" in lib.lbm_last_error()


def test_cell_count_limit():
    """Per-cell indices are 32-bit: 2^32 cells or more are refused."""
    b = ctypes.c_uint64()
    lib = capi.load()
    p = capi.make_params(2048, 2048, 1024)  # exactly 2^32
    assert lib.lbm_bytes_required(ctypes.byref(p), ctypes.byref(b)) == -1
    assert b"2^32" in lib.lbm_last_error()
    p = capi.make_params(2048, 2048, 1023)
    assert lib.lbm_bytes_required(ctypes.byref(p), ctypes.byref(b)) == 0


@pytest.mark.parametrize("args,code,msg", [
    ((9, 8, 8, 4, 4, 256, 7), -1, b"access pattern"),
    ((1, 4, 8, 4, 4, 256, 7), -1, b"elem_bytes"),
    ((1, 8, 0, 4, 4, 256, 7), -1, b"lx, ly, lz"),
    ((1, 8, 8, 4, 4, 64, 7), -5, b"block_x"),
    ((1, 8, 8, 4, 4, 256, 5), -5, b"flags"),
])
def test_pattern_copy_rejects_bad_arguments_without_touching_the_gpu(args, code, msg):
    lib = capi.load()
    buf = ctypes.create_string_buffer(64)
    src, dst = ctypes.c_void_p(ctypes.addressof(buf)), ctypes.c_void_p(ctypes.addressof(buf) + 32)
    assert lib.lbm_pattern_copy(*args, src, dst, None) == code
    assert msg in lib.lbm_last_error()
    v = ctypes.c_double(-1.0)
    assert lib.lbm_pattern_bandwidth(*args, 3, ctypes.byref(v)) == code
    assert v.value == -1.0


def test_pattern_copy_checks_buffers_before_launching():
    """In-place patterns take one buffer, the others two distinct ones; NULL
    buffers and reps < 1 are refused before any device call."""
    lib = capi.load()
    buf = ctypes.create_string_buffer(64)
    a, b = ctypes.c_void_p(ctypes.addressof(buf)), ctypes.c_void_p(ctypes.addressof(buf) + 32)
    shape = (8, 4, 4, 256, 7)
    for pat in (capi.ACCESS_PATTERNS["inplace_even"], capi.ACCESS_PATTERNS["inplace_odd"]):
        assert lib.lbm_pattern_copy(pat, 8, *shape, a, b, None) == -1
        assert b"dst == src" in lib.lbm_last_error()
    for pat in (capi.ACCESS_PATTERNS["copy"], capi.ACCESS_PATTERNS["pull"], capi.ACCESS_PATTERNS["push"]):
        assert lib.lbm_pattern_copy(pat, 2, *shape, a, a, None) == -1
        assert b"dst must differ" in lib.lbm_last_error()
    assert lib.lbm_pattern_copy(1, 8, *shape, None, b, None) == -1
    v = ctypes.c_double()
    assert lib.lbm_pattern_bandwidth(1, 8, *shape, 0, ctypes.byref(v)) == -1
    assert lib.lbm_pattern_bandwidth(1, 8, *shape, 3, None) == -1
