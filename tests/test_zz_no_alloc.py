"""lbm_reduce_fields_noalloc: the diagnostics of lbm_reduce over moment fields
held in device memory by the caller, with every scratch byte in a fixed
device array of the library (no counterpart in the reference; for callers
that cannot allocate).

Checked through the C ABI against lbm_reduce of the same fields, bit for bit
(the same kernels, block width and fixed summation order), on two calls in a
row, the second with a smaller input; the first shape needs more than
8 * 1024 partial records, so it takes the two-level finish, the second
reads fewer records than the first left in the array.  The device buffers
are the test's own (hipMalloc through ctypes): the library allocates nothing.
"""
import ctypes

import numpy as np
import pytest

from lbm_ddf_amd import Cavity, capi

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "ux", "uy", "uz", "phi")
H2D, D2H = 1, 2


def _hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _noalloc_reduce(hip, cav: Cavity, step: int) -> dict:
    """Copies the cavity's fields into the test's own device buffers and runs
    the no-allocation reduction on them."""
    ptrs = []
    try:
        for k in FIELDS:
            host = np.ascontiguousarray(cav.get_field(k))
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), host.nbytes) == 0
            ptrs.append(p)
            assert hip.hipMemcpy(p, host.ctypes.data, host.nbytes, H2D) == 0
        out = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(out), ctypes.sizeof(capi.Diag)) == 0
        ptrs.append(out)
        rc = cav.lib.lbm_reduce_fields_noalloc(ctypes.byref(cav.params), *ptrs[:5], step, out, None)
        assert rc == 0, cav.lib.lbm_last_error()
        assert hip.hipDeviceSynchronize() == 0
        d = capi.Diag()
        assert hip.hipMemcpy(ctypes.addressof(d), out, ctypes.sizeof(d), D2H) == 0
        return d.as_dict()
    finally:
        for p in ptrs:
            hip.hipFree(p)


def test_noalloc_reduce_equals_lbm_reduce_on_two_calls():
    hip = _hip()
    # (64, 96, 96): 9216 records of 64 cells, the two-level finish;
    # (20, 12, 10): 120 records, one level
    for shape, seed, steps in (((64, 96, 96), 3, 25), ((20, 12, 10), 4, 17)):
        print(f"no-alloc reduction: cavity {shape}, {steps} steps", flush=True)
        c = Cavity(*shape, rayleigh=1e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
        c.init_conductive(seed, 1.0e-2)
        c.step(steps)
        want = c.reduce()
        got = _noalloc_reduce(hip, c, want["step"])
        c.close()
        for k, v in want.items():
            assert got[k] == v, (shape, k, got[k], v)
        assert got["nonfinite"] == 0 and got["step"] == steps


def test_noalloc_rejects_shapes_beyond_its_scratch():
    """ceil(64/64) * 512 * 256 = 131072 records exceed the fixed array: an
    error before any launch."""
    lib = capi.load()
    p = capi.make_params(64, 512, 256, rayleigh=1e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    dummy = ctypes.c_void_p(1)  # never dereferenced: the size check comes first
    rc = lib.lbm_reduce_fields_noalloc(ctypes.byref(p), dummy, dummy, dummy, dummy, dummy, 0, dummy, None)
    assert rc == -1
    assert b"scratch" in lib.lbm_last_error()
