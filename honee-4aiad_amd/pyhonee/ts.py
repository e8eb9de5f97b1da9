"""ctypes binding of libhonee_ts.so: the PETSc-free GPU time integrator (explicit RK with the mass KSP, BDF with
Newton-GMRES and point-Jacobi) over a NewtonianProblem.  Loading fails loudly when the library is not built."""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import LIB_DIR
from .driver import CEED_MEM_DEVICE, CEED_MEM_HOST, NewtonianProblem, _ptr, load_libs

TS_RK, TS_BDF, TS_ALPHA, TS_BEULER = 0, 1, 2, 3
PC_JACOBI, PC_PBJACOBI = 0, 1
RK_4, RK_5F, RK_SSP_S3, RK_SSP33, RK_EULER = 0, 1, 2, 3, 4

_lib = {}


class TSOptions(ct.Structure):
    _fields_ = [
        ("type", ct.c_int), ("rk_type", ct.c_int), ("ssp_stages", ct.c_int), ("bdf_order", ct.c_int), ("dt", ct.c_double),
        ("mass_gmres", ct.c_int), ("mass_jacobi_rowsum", ct.c_int), ("mass_rtol", ct.c_double), ("mass_restart", ct.c_int),
        ("mass_max_it", ct.c_int),
        ("snes_rtol", ct.c_double), ("snes_atol", ct.c_double), ("snes_stol", ct.c_double), ("snes_max_it", ct.c_int),
        ("ksp_rtol", ct.c_double), ("ksp_atol", ct.c_double), ("ksp_restart", ct.c_int), ("ksp_max_it", ct.c_int),
        ("ksp_pc", ct.c_int), ("mass_pc", ct.c_int), ("pc_lag", ct.c_int), ("snes_mf_operator", ct.c_int),
        ("alpha_radius", ct.c_double), ("snes_lag_jacobian", ct.c_int), ("snes_lag_jacobian_persists", ct.c_int),
        ("snes_linesearch_type", ct.c_int),
    ]


class TSStats(ct.Structure):
    _fields_ = [
        ("steps", ct.c_int), ("time", ct.c_double),
        ("rhs_evals", ct.c_longlong), ("ifunction_evals", ct.c_longlong), ("jacobian_applies", ct.c_longlong),
        ("diagonal_assemblies", ct.c_longlong), ("mass_applies", ct.c_longlong),
        ("newton_its", ct.c_longlong), ("krylov_its", ct.c_longlong), ("mass_krylov_its", ct.c_longlong),
        ("line_search_backtracks", ct.c_longlong),
        ("last_newton_its", ct.c_int), ("last_krylov_its", ct.c_int), ("last_fnorm0", ct.c_double), ("last_fnorm", ct.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def load_ts_lib():
    if "ts" not in _lib:
        load_libs()
        lib_dir = os.environ.get("HONEE_MI355X_LIB_DIR", LIB_DIR)
        path = os.path.join(lib_dir, "libhonee_ts.so")
        if not os.path.exists(path):
            raise RuntimeError(f"MI355X time integrator library {path} is not built; run __graft_entry__.build()")
        lib = ct.CDLL(path)
        lib.HoneeMiTSOptionsDefault.argtypes = [ct.POINTER(TSOptions)]
        lib.HoneeMiTSOptionsDefault.restype = None
        lib.HoneeMiTSGetTableau.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int)]
        lib.HoneeMiTSCreate.argtypes = [ct.c_void_p, ct.POINTER(TSOptions), ct.POINTER(ct.c_void_p)]
        lib.HoneeMiTSSetSolution.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_double]
        lib.HoneeMiTSGetSolution.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p]
        lib.HoneeMiTSGetSolutionDevice.argtypes = [ct.c_void_p]
        lib.HoneeMiTSGetSolutionDevice.restype = ct.c_void_p
        lib.HoneeMiTSStep.argtypes = [ct.c_void_p, ct.c_int]
        lib.HoneeMiTSGetStats.argtypes = [ct.c_void_p, ct.POINTER(TSStats)]
        lib.HoneeMiTSError.argtypes = [ct.c_void_p]
        lib.HoneeMiTSError.restype = ct.c_char_p
        lib.HoneeMiTSDestroy.argtypes = [ct.POINTER(ct.c_void_p)]
        _lib["ts"] = lib
    return _lib["ts"]


def default_options(**kw) -> TSOptions:
    o = TSOptions()
    load_ts_lib().HoneeMiTSOptionsDefault(ct.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def tableau(rk_type: int, ssp_stages: int = 4):
    """Butcher tableau (A, b, c) of the integrator's explicit schemes (host-side, no GPU needed)."""
    lib = load_ts_lib()
    A, b, c = np.zeros(64 * 64), np.zeros(64), np.zeros(64)
    s = ct.c_int(0)
    ierr = lib.HoneeMiTSGetTableau(int(rk_type), int(ssp_stages), 64, _ptr(A), _ptr(b), _ptr(c), ct.byref(s))
    if ierr:
        raise ValueError(f"no tableau for rk_type={rk_type} ssp_stages={ssp_stages}")
    n = s.value
    return A[: n * n].reshape(n, n).copy(), b[:n].copy(), c[:n].copy()


class TimeStepper:
    """GPU time integrator of a NewtonianProblem; the solution stays in device memory between steps."""

    def __init__(self, prob: NewtonianProblem, options: TSOptions):
        self.lib = load_ts_lib()
        self.prob = prob
        self.options = options
        self.handle = ct.c_void_p()
        ierr = self.lib.HoneeMiTSCreate(prob.handle, ct.byref(options), ct.byref(self.handle))
        if ierr:
            raise RuntimeError(f"HoneeMiTSCreate failed ({ierr}): {self.lib.HoneeMiTSError(None).decode()}")

    def _check(self, ierr, what):
        if ierr:
            raise RuntimeError(f"{what} failed ({ierr}): {self.lib.HoneeMiTSError(self.handle).decode()}")

    def set_solution(self, q: np.ndarray, time: float = 0.0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._check(self.lib.HoneeMiTSSetSolution(self.handle, CEED_MEM_HOST, _ptr(q), float(time)), "SetSolution")

    def solution(self) -> np.ndarray:
        out = np.zeros(5 * self.prob.nnodes)
        self._check(self.lib.HoneeMiTSGetSolution(self.handle, CEED_MEM_HOST, _ptr(out)), "GetSolution")
        return out

    def solution_device_ptr(self) -> int:
        return int(self.lib.HoneeMiTSGetSolutionDevice(self.handle))

    def step(self, nsteps: int = 1):
        self._check(self.lib.HoneeMiTSStep(self.handle, int(nsteps)), "Step")

    def stats(self) -> dict:
        s = TSStats()
        self._check(self.lib.HoneeMiTSGetStats(self.handle, ct.byref(s)), "GetStats")
        return s.as_dict()

    def monitor_ke(self) -> np.ndarray:
        """Kinetic-energy monitor sums of the current (device-resident) solution."""
        out = np.zeros(5)
        self.prob._check(self.prob.lib.HoneeMiMonitorKE(self.prob.handle, CEED_MEM_DEVICE, ct.c_void_p(self.solution_device_ptr()),
                                                        out.ctypes.data_as(ct.POINTER(ct.c_double))), "MonitorKE")
        return out

    def close(self):
        if self.handle:
            self.lib.HoneeMiTSDestroy(ct.byref(self.handle))
            self.handle = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
