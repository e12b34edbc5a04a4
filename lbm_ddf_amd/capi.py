"""ctypes binding of the C ABI in include/lbm_ddf.h.

The product path is the HIP library ``lbm_ddf_amd/_lib/liblbm_ddf.so``; this
module only loads it and marshals host buffers.  There is no fallback: if the
library is missing, loading raises.

``Cavity`` mirrors the reference's stage interface (src/rb3d.h:47-68):
``init_arrays`` (src/initial.cu:208-242), ``forcing`` (src/macrovar.cu:36-47),
``collision_BGK`` / ``collision_BGK_scalar`` (src/collision.cu),
``streaming`` / ``streaming_scalar`` (src/streaming.cu, with the copy-back of
src/main.cu:39,41) and ``macrovar`` (src/macrovar.cu:3-34), plus the fused
``step`` and the diagnostics of src/output.cu.  Errors raise ``LbmError``
instead of exiting the process as the reference's CHECK_CUDA_ERROR does
(src/rb3d.h:71-79).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_lib" / "liblbm_ddf.so"

PREC = {"fp64": 0, "fp16": 1, "bf16": 2}
FIELDS = {"rho": 0, "ux": 1, "uy": 2, "uz": 3, "phi": 4, "fx": 5, "fy": 6, "fz": 7}
POPS = {"f": 0, "g": 1}


class LbmError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("lx", ctypes.c_int32), ("ly", ctypes.c_int32), ("lz", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("rayleigh", ctypes.c_double), ("prandtl", ctypes.c_double),
        ("t_hot", ctypes.c_double), ("t_cold", ctypes.c_double),
        ("gravity", ctypes.c_double), ("beta", ctypes.c_double),
        ("pattern", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


PATTERN = {"ab": 0, "aa": 1}


class Derived(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("visc", "tau", "u0", "diff", "tauc", "t0")]


class Diag(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int64)] + [
        (k, ctypes.c_double)
        for k in ("nu_bulk", "nu_top", "nu_bottom", "umean", "vmean", "wmean", "tmean",
                  "kinetic_energy", "mass", "max_speed", "nonfinite")
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)

# Every entry point declared in include/lbm_ddf.h: name -> (restype, argtypes).
SIGNATURES = {
    "lbm_last_error": (ctypes.c_char_p, []),
    "lbm_derive": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(Derived)]),
    "lbm_bytes_required": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_uint64)]),
    "lbm_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    "lbm_destroy": (ctypes.c_int, [_P]),
    "lbm_init_arrays": (ctypes.c_int, [_P]),
    "lbm_init_conductive": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_double]),
    "lbm_forcing": (ctypes.c_int, [_P]),
    "lbm_collision_BGK": (ctypes.c_int, [_P]),
    "lbm_collision_BGK_scalar": (ctypes.c_int, [_P]),
    "lbm_streaming": (ctypes.c_int, [_P]),
    "lbm_streaming_scalar": (ctypes.c_int, [_P]),
    "lbm_macrovar": (ctypes.c_int, [_P]),
    "lbm_step_staged": (ctypes.c_int, [_P, ctypes.c_int64]),
    "lbm_step": (ctypes.c_int, [_P, ctypes.c_int64]),
    "lbm_step_timed": (ctypes.c_int, [_P, ctypes.c_int64, _D]),
    "lbm_run_diag": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(Diag),
                                    ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "lbm_finalize": (ctypes.c_int, [_P]),
    "lbm_reduce": (ctypes.c_int, [_P, ctypes.POINTER(Diag)]),
    "lbm_profile": (ctypes.c_int, [_P, _D]),
    "lbm_step_count": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "lbm_synchronize": (ctypes.c_int, [_P]),
    "lbm_get_field": (ctypes.c_int, [_P, ctypes.c_int32, _D]),
    "lbm_set_field": (ctypes.c_int, [_P, ctypes.c_int32, _D]),
    "lbm_get_pops": (ctypes.c_int, [_P, ctypes.c_int32, _D]),
    "lbm_set_pops": (ctypes.c_int, [_P, ctypes.c_int32, _D]),
    "lbm_set_tuning": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "lbm_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "lbm_autotune": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), _D]),
    "lbm_copy_bandwidth": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, _D]),
    # pattern, elem_bytes, lx, ly, lz, block_x, flags, src, dst (device pointers as integers), stream
    "lbm_pattern_copy": (ctypes.c_int, [ctypes.c_int32] * 7 + [ctypes.c_void_p] * 3),
    "lbm_pattern_bandwidth": (ctypes.c_int, [ctypes.c_int32] * 8 + [_D]),
    "lbm_pattern_time": (ctypes.c_int, [_P, ctypes.c_int32, _D]),
    "lbm_output_flow": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "lbm_output_fg": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "lbm_restart": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64]),
    "lbm_ensemble_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(_P)]),
    "lbm_ensemble_destroy": (ctypes.c_int, [_P]),
    "lbm_ensemble_add": (ctypes.c_int, [_P, ctypes.POINTER(Params), ctypes.c_uint64, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_int32)]),
    "lbm_ensemble_remove": (ctypes.c_int, [_P, ctypes.c_int32]),
    "lbm_ensemble_count": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "lbm_ensemble_cavity": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(_P)]),
    "lbm_ensemble_step": (ctypes.c_int, [_P, ctypes.c_int64]),
    "lbm_ensemble_step_timed": (ctypes.c_int, [_P, ctypes.c_int64, _D]),
    "lbm_ensemble_run_diag": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(Diag), ctypes.c_int32]),
    "lbm_ensemble_set_tuning": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32]),
    "lbm_ensemble_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                               ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                               ctypes.POINTER(ctypes.c_int32)]),
    "lbm_ensemble_autotune": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_char_p, _D]),
    # device pointers as integers (c_void_p): fields, the output record, the stream
    "lbm_reduce_fields_noalloc": (ctypes.c_int, [ctypes.POINTER(Params)] + [ctypes.c_void_p] * 5 +
                                  [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
}


def load(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load the HIP library (in-tree build).  Raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise LbmError(f"HIP library not built: {p} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load().lbm_last_error()
        raise LbmError(f"lbm error {rc}: {msg.decode() if msg else ''}")


def _dptr(a: np.ndarray):
    # checked always (not an assert): the library writes through this pointer
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_D)


def make_params(lx, ly, lz, rayleigh=1.0e7, prandtl=0.71, t_hot=1.0, t_cold=0.0, gravity=1.0,
                beta=0.0, precision="fp64", pattern="ab") -> Params:
    """Runtime form of the macros of src/parameters.h:8-26 (defaults are the
    reference's; beta <= 0 selects 0.1/(5*LY)).  pattern "ab" keeps two
    population buffers per distribution, "aa" one (updated in place)."""
    return Params(int(lx), int(ly), int(lz), PREC[precision], float(rayleigh), float(prandtl),
                  float(t_hot), float(t_cold), float(gravity), float(beta), PATTERN[pattern], 0)


def derive(params: Params) -> dict:
    """src/globals.cu:11-23."""
    d = Derived()
    _check(load().lbm_derive(ctypes.byref(params), ctypes.byref(d)))
    return {k: getattr(d, k) for k, _ in d._fields_}


def copy_bandwidth(nbytes: int = 4 << 30, reps: int = 10) -> float:
    """GB/s (read + write) of a plain device copy kernel."""
    v = ctypes.c_double()
    _check(load().lbm_copy_bandwidth(ctypes.c_uint64(int(nbytes)), int(reps), ctypes.byref(v)))
    return v.value


# Access patterns of lbm_pattern_copy (include/lbm_ddf.h LBM_ACCESS_*).
ACCESS_PATTERNS = {"copy": 0, "pull": 1, "push": 2, "inplace_even": 3, "inplace_odd": 4}


def pattern_copy(pattern: str, elem_bytes: int, shape: tuple, src_ptr: int, dst_ptr: int, block_x: int = 256,
                 flags: int = 7, stream: int = 0) -> None:
    """Enqueue one access-pattern copy over device buffers of 54*lx*ly*lz elements."""
    lx, ly, lz = shape
    _check(load().lbm_pattern_copy(ACCESS_PATTERNS[pattern], int(elem_bytes), int(lx), int(ly), int(lz), int(block_x),
                                   int(flags), ctypes.c_void_p(src_ptr), ctypes.c_void_p(dst_ptr),
                                   ctypes.c_void_p(stream or None)))


def pattern_bandwidth(pattern: str, elem_bytes: int, shape: tuple, block_x: int = 256, flags: int = 7,
                      reps: int = 10) -> float:
    """Mean milliseconds of one access-pattern copy pass (own buffers)."""
    lx, ly, lz = shape
    v = ctypes.c_double()
    _check(load().lbm_pattern_bandwidth(ACCESS_PATTERNS[pattern], int(elem_bytes), int(lx), int(ly), int(lz),
                                        int(block_x), int(flags), int(reps), ctypes.byref(v)))
    return v.value


def bytes_required(params: Params) -> int:
    v = ctypes.c_uint64()
    _check(load().lbm_bytes_required(ctypes.byref(params), ctypes.byref(v)))
    return int(v.value)


class Cavity:
    """One Rayleigh-Benard cavity on the current HIP device."""

    def __init__(self, lx, ly, lz, **kw):
        self.lib = load()
        self.params = make_params(lx, ly, lz, **kw)
        self.shape = (int(lz), int(ly), int(lx))
        self.n = int(lx) * int(ly) * int(lz)
        h = _P()
        _check(self.lib.lbm_create(ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- initial state ----------------------------------------------------
    def init_arrays(self):
        _check(self.lib.lbm_init_arrays(self.h))

    def init_conductive(self, seed: int, amp: float = 1.0e-3):
        _check(self.lib.lbm_init_conductive(self.h, ctypes.c_uint64(seed), float(amp)))

    # -- stage interface --------------------------------------------------
    def forcing(self):
        _check(self.lib.lbm_forcing(self.h))

    def collision_BGK(self):
        _check(self.lib.lbm_collision_BGK(self.h))

    def collision_BGK_scalar(self):
        _check(self.lib.lbm_collision_BGK_scalar(self.h))

    def streaming(self):
        _check(self.lib.lbm_streaming(self.h))

    def streaming_scalar(self):
        _check(self.lib.lbm_streaming_scalar(self.h))

    def macrovar(self):
        _check(self.lib.lbm_macrovar(self.h))

    def staged_step(self, n: int = 1):
        """n iterations of src/main.cu:27-48 through the stage entry points
        (from Python, one call per stage)."""
        for _ in range(n):
            self.forcing()
            self.collision_BGK()
            self.collision_BGK_scalar()
            self.streaming()
            self.streaming_scalar()
            self.macrovar()

    def step_staged(self, n: int = 1):
        """lbm_step_staged: the reference's loop (six launches, three host
        synchronizations per step) driven from C++."""
        _check(self.lib.lbm_step_staged(self.h, int(n)))

    # -- fused path -------------------------------------------------------
    def step(self, n: int = 1):
        _check(self.lib.lbm_step(self.h, int(n)))

    def step_timed(self, n: int) -> float:
        ms = ctypes.c_double()
        _check(self.lib.lbm_step_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def pattern_time(self, reps: int = 10) -> tuple:
        """(ms, ms) of the access-pattern copy of this context's fused pass
        on its own buffers (A/B: pull 0->1 and 1->0; in place: even, odd).
        Overwrites the populations: initialise again before stepping."""
        ms = (ctypes.c_double * 2)()
        _check(self.lib.lbm_pattern_time(self.h, int(reps), ms))
        return ms[0], ms[1]

    def run_diag(self, n: int, every: int) -> list:
        cap = n // every + 1
        arr = (Diag * cap)()
        nout = ctypes.c_int64()
        _check(self.lib.lbm_run_diag(self.h, int(n), int(every), arr, cap, ctypes.byref(nout)))
        return [arr[i].as_dict() for i in range(nout.value)]

    def set_tuning(self, block_x: int, flags: int = 0):
        _check(self.lib.lbm_set_tuning(self.h, int(block_x), int(flags)))

    @property
    def tuning(self) -> tuple:
        bx, fl = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.lbm_get_tuning(self.h, ctypes.byref(bx), ctypes.byref(fl)))
        return int(bx.value), int(fl.value)

    def autotune(self, steps: int = 3) -> dict:
        taken, best = ctypes.c_int64(), ctypes.c_double()
        _check(self.lib.lbm_autotune(self.h, int(steps), ctypes.byref(taken), ctypes.byref(best)))
        bx, fl = self.tuning
        return {"block_x": bx, "flags": fl, "steps_taken": int(taken.value), "ms_per_step": best.value}

    def finalize(self):
        _check(self.lib.lbm_finalize(self.h))

    def synchronize(self):
        _check(self.lib.lbm_synchronize(self.h))

    @property
    def step_count(self) -> int:
        v = ctypes.c_int64()
        _check(self.lib.lbm_step_count(self.h, ctypes.byref(v)))
        return int(v.value)

    # -- diagnostics ------------------------------------------------------
    def reduce(self) -> dict:
        d = Diag()
        _check(self.lib.lbm_reduce(self.h, ctypes.byref(d)))
        return d.as_dict()

    def profile(self) -> np.ndarray:
        out = np.zeros((self.shape[1], 5), dtype=np.float64)
        _check(self.lib.lbm_profile(self.h, _dptr(out)))
        return out

    # -- host copies ------------------------------------------------------
    def get_field(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(self.lib.lbm_get_field(self.h, FIELDS[name], _dptr(out)))
        return out

    def set_field(self, name: str, values) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), self.shape))
        _check(self.lib.lbm_set_field(self.h, FIELDS[name], _dptr(a)))

    def get_pops(self, which: str) -> np.ndarray:
        out = np.empty((27,) + self.shape, dtype=np.float64)
        _check(self.lib.lbm_get_pops(self.h, POPS[which], _dptr(out)))
        return out

    def set_pops(self, which: str, values) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), (27,) + self.shape))
        _check(self.lib.lbm_set_pops(self.h, POPS[which], _dptr(a)))

    # -- checkpoints (reference binary formats) ----------------------------
    def output_flow(self, path) -> None:
        """output_flow (src/output.cu:67-146): rho', ux, uy, uz, phi as FP64."""
        _check(self.lib.lbm_output_flow(self.h, os.fsencode(path)))

    def output_fg(self, path) -> None:
        """output_fg (src/output.cu:379-419): f then g as FP64."""
        _check(self.lib.lbm_output_fg(self.h, os.fsencode(path)))

    def restart(self, flow_path, fg_path, step: int) -> None:
        """initialize() with CONTINUE_STEP (src/initial.cu:110-205)."""
        _check(self.lib.lbm_restart(self.h, os.fsencode(flow_path), os.fsencode(fg_path), int(step)))

    def state(self) -> dict:
        """The reference state: f, g, rho', u, phi."""
        return {
            "f": self.get_pops("f"), "g": self.get_pops("g"),
            "rho": self.get_field("rho"), "ux": self.get_field("ux"), "uy": self.get_field("uy"),
            "uz": self.get_field("uz"), "phi": self.get_field("phi"),
        }


class _Member(Cavity):
    """A cavity owned by an Ensemble: read through the single-cavity calls."""

    def __init__(self, ens, slot, params):
        self.lib = load()
        self.params = params
        self.shape = (int(params.lz), int(params.ly), int(params.lx))
        self.n = int(params.lx) * int(params.ly) * int(params.lz)
        self._ens = ens
        self.slot = slot
        h = _P()
        _check(self.lib.lbm_ensemble_cavity(ens.h, slot, ctypes.byref(h)))
        self.h = h

    def close(self):  # owned by the ensemble
        self.h = None


class Ensemble:
    """Many cavities advanced by one launch per step (lbm_ensemble_*)."""

    def __init__(self, precision: str = "fp64"):
        self.lib = load()
        self.precision = precision
        h = _P()
        _check(self.lib.lbm_ensemble_create(PREC[precision], ctypes.byref(h)))
        self.h = h
        self.params = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_ensemble_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, lx, ly, lz, seed: int, amp: float = 1.0e-3, **kw) -> int:
        p = make_params(lx, ly, lz, precision=self.precision, **kw)
        slot = ctypes.c_int32()
        _check(self.lib.lbm_ensemble_add(self.h, ctypes.byref(p), ctypes.c_uint64(seed), float(amp),
                                         ctypes.byref(slot)))
        self.params[slot.value] = p
        return slot.value

    def remove(self, slot: int):
        _check(self.lib.lbm_ensemble_remove(self.h, int(slot)))
        self.params.pop(slot, None)

    def count(self) -> dict:
        s, r, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_uint64()
        _check(self.lib.lbm_ensemble_count(self.h, ctypes.byref(s), ctypes.byref(r), ctypes.byref(n)))
        return {"slots": s.value, "resident": r.value, "cells": int(n.value)}

    def cavity(self, slot: int) -> Cavity:
        return _Member(self, slot, self.params[slot])

    def step(self, n: int):
        _check(self.lib.lbm_ensemble_step(self.h, int(n)))

    def step_timed(self, n: int) -> float:
        ms = ctypes.c_double()
        _check(self.lib.lbm_ensemble_step_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def run_diag(self, n: int) -> dict:
        slots = self.count()["slots"]
        arr = (Diag * max(1, slots))()
        _check(self.lib.lbm_ensemble_run_diag(self.h, int(n), arr, max(1, slots)))
        return {s: arr[s].as_dict() for s in self.params}

    def set_tuning(self, block_x: int, blocks_per_cu: int, graph_pairs: int, flags: int = 7,
                   cells_per_thread: int = 1):
        _check(self.lib.lbm_ensemble_set_tuning(self.h, int(block_x), int(blocks_per_cu), int(graph_pairs),
                                                int(flags), int(cells_per_thread)))

    @property
    def tuning(self) -> tuple:
        """(block_x, blocks_per_cu, graph_pairs, flags, cells_per_thread)."""
        v = [ctypes.c_int32() for _ in range(5)]
        _check(self.lib.lbm_ensemble_get_tuning(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, steps: int = 8, cache_path: str | None = None) -> float:
        ms = ctypes.c_double()
        path = cache_path.encode() if cache_path else None
        _check(self.lib.lbm_ensemble_autotune(self.h, int(steps), path, ctypes.byref(ms)))
        return ms.value


# ---------------------------------------------------------------- serving
class Request(ctypes.Structure):
    _fields_ = [("lx", ctypes.c_int32), ("ly", ctypes.c_int32), ("lz", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("rayleigh", ctypes.c_double), ("prandtl", ctypes.c_double),
                ("t_hot", ctypes.c_double), ("t_cold", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("amp", ctypes.c_double), ("arrival_s", ctypes.c_double), ("max_steps", ctypes.c_int64),
                ("nu_tol", ctypes.c_double), ("window", ctypes.c_int64)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("reserved", ctypes.c_int32), ("arrival_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double), ("steps", ctypes.c_int64),
                ("nu_bulk", ctypes.c_double), ("nu_top", ctypes.c_double), ("nu_bottom", ctypes.c_double),
                ("kinetic_energy", ctypes.c_double), ("nu_running", ctypes.c_double)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}
        d["status"] = REQ_STATUS[self.status]
        return d


class ServeOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("max_resident", ctypes.c_int32), ("chunk_steps", ctypes.c_int64),
                ("hbm_budget", ctypes.c_uint64), ("time_limit_s", ctypes.c_double), ("block_x", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("flags", ctypes.c_int32), ("cells_per_thread", ctypes.c_int32),
                ("max_resident_cells", ctypes.c_uint64), ("progress_s", ctypes.c_double),
                ("pattern", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("busy_s", ctypes.c_double), ("cell_updates", ctypes.c_double),
                ("mlups", ctypes.c_double), ("completed", ctypes.c_int32), ("rejected", ctypes.c_int32),
                ("diverged", ctypes.c_int32), ("unfinished", ctypes.c_int32), ("pending", ctypes.c_int32),
                ("peak_resident", ctypes.c_int32), ("peak_bytes", ctypes.c_uint64), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("arrived", ctypes.c_int32), ("deferred", ctypes.c_int32),
                ("p50_all_s", ctypes.c_double), ("p99_all_s", ctypes.c_double), ("mean_queue_s", ctypes.c_double),
                ("mean_service_s", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_STATUS = {0: "converged", 1: "budget", 2: "rejected", 3: "diverged", 4: "unfinished", 5: "pending"}
SIGNATURES["lbm_serve"] = (ctypes.c_int, [ctypes.POINTER(Request), ctypes.c_int32, ctypes.POINTER(ServeOptions),
                                          ctypes.POINTER(Result), ctypes.POINTER(ServeStats)])


def serve(requests: list, precision: str = "fp64", chunk_steps: int = 100, hbm_budget: float = 250e9,
          time_limit_s: float = 60.0, max_resident: int = 0, block_x: int = 0, blocks_per_cu: int = 0,
          flags: int = 7, max_resident_cells: int = 0, cells_per_thread: int = 1, progress_s: float = 0.0,
          pattern: str = "ab"):
    """Run lbm_serve over a list of request dicts (keys of Request; defaults:
    walls +-0.5, amp 1e-3).  Returns (results list of dicts, stats dict).
    progress_s > 0 makes the loop print a progress line to stderr at that
    interval; pattern "aa" keeps every resident cavity in place (half the
    population bytes; FP64 results equal to "ab").  block_x 0 keeps the
    ensemble's default launch (one block per tile); block_x > 0 sets
    (block_x, blocks_per_cu, flags, cells_per_thread), blocks_per_cu 0
    meaning one block per tile."""
    lib = load()
    n = len(requests)
    arr = (Request * max(1, n))()
    for i, r in enumerate(requests):
        q = dict(t_hot=0.5, t_cold=-0.5, amp=1.0e-3, prandtl=0.71, reserved=0)
        q.update(r)
        for k, _ in Request._fields_:
            setattr(arr[i], k, q[k])
    opt = ServeOptions(PREC[precision], int(max_resident), int(chunk_steps), int(hbm_budget), float(time_limit_s),
                       int(block_x), int(blocks_per_cu), int(flags), int(cells_per_thread), int(max_resident_cells),
                       float(progress_s), PATTERN[pattern], 0)
    res = (Result * max(1, n))()
    st = ServeStats()
    _check(lib.lbm_serve(arr, n, ctypes.byref(opt), res, ctypes.byref(st)))
    return [res[i].as_dict() for i in range(n)], st.as_dict()
