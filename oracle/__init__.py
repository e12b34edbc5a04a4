"""ctypes binding of the FP64 CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: the expected results of the parity tests and the
CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline legs (timed beside the GPU, and in the 64^3 parity workload compared
with it) may import this package; the product (lbm_ddf_amd) never does.

Arrays are numpy-owned, shaped like the reference's: populations
(27, LZ, LY, LX) and fields (LZ, LY, LX), x fastest.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "liboracle.so"


class OrcParams(ctypes.Structure):
    _fields_ = [("lx", ctypes.c_int), ("ly", ctypes.c_int), ("lz", ctypes.c_int)] + [
        (k, ctypes.c_double)
        for k in ("rayl", "prand", "t_hot", "t_cold", "grav0", "beta",
                  "visc", "tau", "u0", "diff", "tauc", "t0")
    ]


_D = ctypes.POINTER(ctypes.c_double)


class OrcFields(ctypes.Structure):
    _fields_ = [(k, _D) for k in ("f", "g", "f_tmp", "g_tmp", "rho", "ux", "uy", "uz", "phi",
                                  "fx", "fy", "fz")]


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (make -C oracle)")
        lib = ctypes.CDLL(str(LIB_PATH))
        P, F = ctypes.POINTER(OrcParams), ctypes.POINTER(OrcFields)
        lib.orc_tables.argtypes = [ctypes.POINTER(ctypes.c_int)] * 3 + [_D]
        lib.orc_derive.argtypes = [P]
        for name in ("orc_init_arrays", "orc_forcing", "orc_collision_bgk", "orc_collision_bgk_scalar",
                     "orc_streaming", "orc_streaming_scalar", "orc_macrovar"):
            getattr(lib, name).argtypes = [P, F]
        lib.orc_init_conductive.argtypes = [P, F, ctypes.c_uint64, ctypes.c_double]
        lib.orc_perturbation.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double]
        lib.orc_perturbation.restype = ctypes.c_double
        lib.orc_run.argtypes = [P, F, ctypes.c_long]
        for name in ("orc_output_nu", "orc_diag_flow", "orc_output_profile", "orc_energy_mass"):
            getattr(lib, name).argtypes = [P, F, _D]
        _lib = lib
    return _lib


def tables():
    """(cix, ciy, ciz, tp) of src/initial.cu:71-83."""
    lib = load()
    c = [(ctypes.c_int * 27)() for _ in range(3)]
    tp = (ctypes.c_double * 27)()
    lib.orc_tables(c[0], c[1], c[2], tp)
    return (np.array(c[0][:]), np.array(c[1][:]), np.array(c[2][:]), np.array(tp[:]))


def make_params(lx, ly, lz, rayleigh=1.0e7, prandtl=0.71, t_hot=1.0, t_cold=0.0, gravity=1.0,
                beta=0.0) -> OrcParams:
    p = OrcParams()
    p.lx, p.ly, p.lz = int(lx), int(ly), int(lz)
    p.rayl, p.prand, p.t_hot, p.t_cold = float(rayleigh), float(prandtl), float(t_hot), float(t_cold)
    p.grav0, p.beta = float(gravity), float(beta)
    load().orc_derive(ctypes.byref(p))
    return p


def perturbation(seed: int, idx: int, amp: float) -> float:
    return load().orc_perturbation(ctypes.c_uint64(seed), ctypes.c_uint64(idx), float(amp))


class OracleState:
    """Host state of one cavity plus the oracle stage functions."""

    POP_NAMES = ("f", "g", "f_tmp", "g_tmp")
    FIELD_NAMES = ("rho", "ux", "uy", "uz", "phi", "fx", "fy", "fz")

    def __init__(self, lx, ly, lz, **kw):
        self.p = make_params(lx, ly, lz, **kw)
        self.shape = (int(lz), int(ly), int(lx))
        self.arrays = {k: np.zeros((27,) + self.shape) for k in self.POP_NAMES}
        self.arrays.update({k: np.zeros(self.shape) for k in self.FIELD_NAMES})
        self._rebind()

    def _rebind(self):
        self.fields = OrcFields(*[self.arrays[k].ctypes.data_as(_D) for k, _ in OrcFields._fields_])

    def __getitem__(self, k):
        return self.arrays[k]

    def __setitem__(self, k, v):
        self.arrays[k][...] = v

    def _call(self, name, *extra):
        getattr(load(), name)(ctypes.byref(self.p), ctypes.byref(self.fields), *extra)

    # src/ stage functions
    def init_arrays(self):
        self._call("orc_init_arrays")

    def init_conductive(self, seed: int, amp: float = 1.0e-3):
        self._call("orc_init_conductive", ctypes.c_uint64(seed), float(amp))

    def forcing(self):
        self._call("orc_forcing")

    def collision_BGK(self):
        self._call("orc_collision_bgk")

    def collision_BGK_scalar(self):
        self._call("orc_collision_bgk_scalar")

    def streaming(self):
        self._call("orc_streaming")

    def streaming_scalar(self):
        self._call("orc_streaming_scalar")

    def macrovar(self):
        self._call("orc_macrovar")

    def run(self, nsteps: int):
        self._call("orc_run", ctypes.c_long(int(nsteps)))

    # diagnostics
    def output_nu(self):
        out = (ctypes.c_double * 3)()
        self._call("orc_output_nu", out)
        return tuple(out)

    def diag_flow(self):
        out = (ctypes.c_double * 4)()
        self._call("orc_diag_flow", out)
        return tuple(out)

    def profile(self):
        out = np.zeros((self.shape[1], 5))
        self._call("orc_output_profile", out.ctypes.data_as(_D))
        return out

    def energy_mass(self):
        out = (ctypes.c_double * 2)()
        self._call("orc_energy_mass", out)
        return tuple(out)

    def diag(self) -> dict:
        nu = self.output_nu()
        m = self.diag_flow()
        ke, mass = self.energy_mass()
        return {"nu_bulk": nu[0], "nu_top": nu[1], "nu_bottom": nu[2], "umean": m[0], "vmean": m[1],
                "wmean": m[2], "tmean": m[3], "kinetic_energy": ke, "mass": mass}
