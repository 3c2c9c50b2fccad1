"""MI355X-native explicit residual path of CTurboBFM (JST / Roe-MUSCL fluxes, Hall /
lift-drag body forces, local / global time step, RK41A / RK31B).

The product is the native library ``libtbfm_gpu.so`` (HIP kernels for gfx950 behind the
C-ABI of ``include/tbfm_gpu.h``).  This module binds it with ctypes and offers a small
host-side mirror of the reference's ``Solver`` surface (src/solver.cpp, include/solver.hpp):

    s = Solver("input.ini")          # or Solver.from_case(cases.rotor3d(...))
    s.solve()                        # N_ITERATIONS or until the convergence test passes

There is no CPU fallback: without the built library every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

from . import cases  # noqa: F401  (synthetic workloads)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtbfm_gpu.so")

TBFM_OK = 0
TBFM_ERR_DIVERGED = -4


class CaseDesc(C.Structure):
    _fields_ = [("ni", C.c_long), ("nj", C.c_long), ("nk", C.c_long),
                ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
                ("nfields", C.c_int), ("field_names", C.POINTER(C.c_char_p)),
                ("field_data", C.POINTER(C.POINTER(C.c_double))),
                ("nconfig", C.c_int), ("config_keys", C.POINTER(C.c_char_p)),
                ("config_values", C.POINTER(C.c_char_p)),
                ("npatches", C.c_int), ("patch_names", C.POINTER(C.c_char_p)),
                ("patch_ranges", C.POINTER(C.c_long))]


EXPORTS = ["tbfm_gpu_case_create", "tbfm_gpu_case_create_from_ini", "tbfm_gpu_case_destroy",
           "tbfm_gpu_last_error", "tbfm_gpu_dims", "tbfm_gpu_set_state", "tbfm_gpu_get_state",
           "tbfm_gpu_advance", "tbfm_gpu_compute_residual", "tbfm_gpu_compute_timestep",
           "tbfm_gpu_get_side_outputs", "tbfm_gpu_get_turbo", "tbfm_gpu_get_geometry",
           "tbfm_gpu_profile", "tbfm_gpu_profile_get", "tbfm_gpu_profile_get_prestage", "tbfm_gpu_sync", "tbfm_gpu_copy_calibration",
           "tbfm_gpu_tuning", "tbfm_gpu_case_device_bytes", "tbfm_gpu_case_set_convergence",
           "tbfm_gpu_case_status", "tbfm_gpu_batch_create", "tbfm_gpu_batch_admit", "tbfm_gpu_batch_retire",
           "tbfm_gpu_batch_size", "tbfm_gpu_batch_advance", "tbfm_gpu_batch_destroy",
           "tbfm_gpu_batch_last_error", "tbfm_gpu_device_memory", "tbfm_gpu_write_solution",
           "tbfm_write_solution_from_ini", "tbfm_volume_file_name", "tbfm_initial_state_from_ini",
           "tbfm_gpu_get_turbo_history", "tbfm_gpu_compute_timestep_noalloc"]

_lib = None


class SolverError(RuntimeError):
    pass


def lib():
    """The native library; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SolverError("native library missing: %s -- run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.tbfm_gpu_case_create.argtypes = [C.POINTER(CaseDesc), C.POINTER(vp)]
        L.tbfm_gpu_case_create_from_ini.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.tbfm_gpu_case_destroy.argtypes = [vp]
        L.tbfm_gpu_last_error.restype = C.c_char_p
        L.tbfm_gpu_last_error.argtypes = [vp]
        L.tbfm_gpu_dims.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.tbfm_gpu_set_state.argtypes = [vp, dp]
        L.tbfm_gpu_get_state.argtypes = [vp, dp]
        L.tbfm_gpu_advance.argtypes = [vp, C.c_long, dp, C.POINTER(C.c_long)]
        L.tbfm_gpu_compute_residual.argtypes = [vp, dp, C.c_long, C.c_int, dp]
        L.tbfm_gpu_compute_timestep.argtypes = [vp, dp, dp]
        L.tbfm_gpu_compute_timestep_noalloc.argtypes = [vp, vp, vp, vp]
        L.tbfm_gpu_get_side_outputs.argtypes = [vp, dp, dp, dp, dp]
        L.tbfm_gpu_get_turbo.argtypes = [vp, dp]
        L.tbfm_gpu_get_geometry.argtypes = [vp, C.c_int, dp]
        L.tbfm_gpu_profile.argtypes = [vp, C.c_int, C.c_int]
        L.tbfm_gpu_profile_get.argtypes = [vp, dp, C.POINTER(C.c_long), dp]
        L.tbfm_gpu_profile_get_prestage.argtypes = [vp, dp, dp]
        L.tbfm_gpu_sync.argtypes = [vp]
        L.tbfm_gpu_copy_calibration.argtypes = [C.c_long, C.c_int, dp]
        L.tbfm_gpu_tuning.restype = C.c_char_p
        L.tbfm_gpu_tuning.argtypes = [vp]
        L.tbfm_gpu_case_device_bytes.argtypes = [vp, C.POINTER(C.c_long)]
        L.tbfm_gpu_case_set_convergence.argtypes = [vp, C.c_int, C.c_double]
        L.tbfm_gpu_case_status.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_int), dp]
        L.tbfm_gpu_batch_create.argtypes = [C.POINTER(vp)]
        L.tbfm_gpu_batch_admit.argtypes = [vp, vp]
        L.tbfm_gpu_batch_retire.argtypes = [vp, vp]
        L.tbfm_gpu_batch_size.argtypes = [vp]
        L.tbfm_gpu_batch_advance.argtypes = [vp, C.c_long]
        L.tbfm_gpu_batch_destroy.argtypes = [vp]
        L.tbfm_gpu_batch_last_error.restype = C.c_char_p
        L.tbfm_gpu_batch_last_error.argtypes = [vp]
        L.tbfm_gpu_device_memory.argtypes = [C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.tbfm_gpu_write_solution.argtypes = [vp, C.c_char_p]
        L.tbfm_gpu_get_turbo_history.argtypes = [vp, C.c_long, C.c_long, dp]
        L.tbfm_write_solution_from_ini.argtypes = [C.c_char_p, dp, dp, dp, dp, C.c_char_p]
        L.tbfm_volume_file_name.argtypes = [C.c_char_p, C.c_long, C.c_char_p, C.c_long]
        L.tbfm_initial_state_from_ini.argtypes = [C.c_char_p, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def copy_calibration(nbytes: int, repeats: int = 10) -> float:
    """Milliseconds per plain device copy of nbytes (k_copy_calibration)."""
    ms = C.c_double()
    rc = lib().tbfm_gpu_copy_calibration(nbytes, repeats, C.byref(ms))
    if rc != TBFM_OK:
        raise SolverError(lib().tbfm_gpu_last_error(None).decode())
    return ms.value


def device_memory():
    """(free, total) bytes of device memory."""
    f, t = C.c_long(), C.c_long()
    rc = lib().tbfm_gpu_device_memory(C.byref(f), C.byref(t))
    if rc != TBFM_OK:
        raise SolverError(lib().tbfm_gpu_last_error(None).decode())
    return f.value, t.value


def _host_error():
    return SolverError(lib().tbfm_gpu_last_error(None).decode())


def _grid_dims(ini_path: str):
    """NI, NJ, NK of the grid file an input.ini names (header lines of the grid CSV)."""
    from .cases import read_ini
    cfg = read_ini(ini_path)
    grid = cfg["GRID_FILE"]
    if not os.path.isabs(grid):
        grid = os.path.join(os.path.dirname(os.path.abspath(ini_path)), grid)
    dims = {}
    with open(grid) as f:
        for line in f:
            for key in ("NI", "NJ", "NK"):
                if line.startswith(key + "="):
                    dims[key] = int(line.split("=")[1])
            if len(dims) == 3:
                break
    return dims["NI"], dims["NJ"], dims["NK"]


def initial_state_from_ini(ini_path: str):
    """(U, finv, fvisc) a case created from input.ini starts from: the uniform INIT_*
    state or the restart file (``tbfm_initial_state_from_ini``; host only)."""
    ni, nj, nk = _grid_dims(ini_path)
    U = np.zeros((5, ni, nj, nk))
    fi = np.zeros((ni, nj, nk, 3))
    fv = np.zeros((ni, nj, nk, 3))
    if lib().tbfm_initial_state_from_ini(os.fsencode(ini_path), _dp(U), _dp(fi), _dp(fv)) != TBFM_OK:
        raise _host_error()
    return U, fi, fv


def write_solution_from_ini(ini_path: str, U: np.ndarray, path: str, finv=None, fvisc=None, dev=None) -> None:
    """The reference's volume solution file of a given state (``tbfm_write_solution_from_ini``;
    host only)."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (U, finv, fvisc, dev)]
    ptrs = [_dp(a) if a is not None else None for a in keep]
    if lib().tbfm_write_solution_from_ini(os.fsencode(ini_path), ptrs[0], ptrs[1], ptrs[2], ptrs[3],
                                          os.fsencode(path)) != TBFM_OK:
        raise _host_error()


def volume_file_name(ini_path: str, iteration: int) -> str:
    """Volume_CSV file name (without folder and extension) at an iteration."""
    buf = C.create_string_buffer(4096)
    if lib().tbfm_volume_file_name(os.fsencode(ini_path), iteration, buf, len(buf)) != TBFM_OK:
        raise _host_error()
    return buf.value.decode()


class Solver:
    """One case resident on the GPU (mirrors Solver::Solver + Solver::solve)."""

    def __init__(self, ini_path: Optional[str] = None, _handle=None):
        self.h = None
        if _handle is None:
            h = C.c_void_p()
            rc = lib().tbfm_gpu_case_create_from_ini(os.fsencode(ini_path), C.byref(h))
            if rc != TBFM_OK:
                raise SolverError(lib().tbfm_gpu_last_error(None).decode())
            _handle = h
        self.h = _handle
        ni, nj, nk = C.c_long(), C.c_long(), C.c_long()
        lib().tbfm_gpu_dims(self.h, C.byref(ni), C.byref(nj), C.byref(nk))
        self.ni, self.nj, self.nk = ni.value, nj.value, nk.value
        self.N = self.ni * self.nj * self.nk
        self.iterations = 0
        self.converged_at = 0
        self.history = np.zeros((0, 5))

    @classmethod
    def from_case(cls, case: Dict, overrides: Optional[Dict[str, str]] = None) -> "Solver":
        cfg = dict(case["config"])
        if overrides:
            cfg.update(overrides)
        keep = []

        def arr(a):
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            keep.append(a)
            return _dp(a)

        d = CaseDesc()
        d.ni, d.nj, d.nk = case["ni"], case["nj"], case["nk"]
        d.x, d.y, d.z = arr(case["x"]), arr(case["y"]), arr(case["z"])
        names = list(case["fields"].keys())
        d.nfields = len(names)
        d.field_names = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
        d.field_data = (C.POINTER(C.c_double) * max(len(names), 1))(*[arr(case["fields"][n]) for n in names])
        keys = list(cfg.keys())
        d.nconfig = len(keys)
        d.config_keys = (C.c_char_p * len(keys))(*[k.encode() for k in keys])
        d.config_values = (C.c_char_p * len(keys))(*[str(cfg[k]).encode() for k in keys])
        d.npatches = len(case["patches"])
        d.patch_names = (C.c_char_p * d.npatches)(*[p[0].encode() for p in case["patches"]])
        ranges = np.array([p[1:7] for p in case["patches"]], dtype=np.int64).reshape(-1)
        keep.append(ranges)
        d.patch_ranges = ranges.ctypes.data_as(C.POINTER(C.c_long))
        h = C.c_void_p()
        rc = lib().tbfm_gpu_case_create(C.byref(d), C.byref(h))
        if rc != TBFM_OK:
            raise SolverError(lib().tbfm_gpu_last_error(None).decode())
        return cls(_handle=h)

    def close(self):
        if self.h is not None and _lib is not None:
            _lib.tbfm_gpu_case_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != TBFM_OK:
            raise SolverError(lib().tbfm_gpu_last_error(self.h).decode())

    # -- state -------------------------------------------------------------------
    def state(self) -> np.ndarray:
        U = np.empty((5, self.ni, self.nj, self.nk))
        self._check(lib().tbfm_gpu_get_state(self.h, _dp(U)))
        return U

    def set_state(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, dtype=np.float64)
        self._check(lib().tbfm_gpu_set_state(self.h, _dp(U)))

    # -- iterations --------------------------------------------------------------
    def advance(self, n: int) -> np.ndarray:
        """Runs up to n iterations; returns their log10 residual norms (rows stop at
        convergence)."""
        lr = np.zeros((n, 5))
        conv = C.c_long()
        rc = lib().tbfm_gpu_advance(self.h, n, _dp(lr), C.byref(conv))
        first = self.iterations
        rows = n if conv.value == 0 or conv.value <= first else conv.value - first
        self.iterations += rows
        self.converged_at = conv.value
        self.history = np.vstack([self.history, lr[:rows]])
        self._check(rc)
        return lr[:rows]

    def solve(self, n_max: Optional[int] = None, chunk: int = 500) -> np.ndarray:
        n_max = n_max if n_max is not None else 10 ** 9
        while self.iterations < n_max and self.converged_at == 0:
            self.advance(min(chunk, n_max - self.iterations))
        return self.history

    def residual(self, U: np.ndarray, it: int = 1, update_profiles: bool = True) -> np.ndarray:
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros_like(U)
        self._check(lib().tbfm_gpu_compute_residual(self.h, _dp(U), it, int(update_profiles), _dp(R)))
        return R

    def timestep(self, U: np.ndarray) -> np.ndarray:
        U = np.ascontiguousarray(U, dtype=np.float64)
        dt = np.zeros((self.ni, self.nj, self.nk))
        self._check(lib().tbfm_gpu_compute_timestep(self.h, _dp(U), _dp(dt)))
        return dt

    def side_outputs(self):
        finv = np.zeros((self.ni, self.nj, self.nk, 3))
        fvisc = np.zeros_like(finv)
        dev = np.zeros((self.ni, self.nj, self.nk))
        dt = np.zeros_like(dev)
        self._check(lib().tbfm_gpu_get_side_outputs(self.h, _dp(finv), _dp(fvisc), _dp(dev), _dp(dt)))
        return finv, fvisc, dev, dt

    def write_solution(self, path: str) -> None:
        """The reference's volume solution file (Output::writeSolution) of the current
        state and side outputs."""
        self._check(lib().tbfm_gpu_write_solution(self.h, os.fsencode(path)))

    def turbo_history(self, first: int = 1, n: Optional[int] = None) -> np.ndarray:
        """Per-iteration rows (time after the iteration, mass flow, PRtt, TRtt, ETAtt) of
        a case created with TURBO_OUTPUT = yes."""
        if n is None:
            n = self.iterations - first + 1
        out = np.zeros((max(n, 0), 5))
        self._check(lib().tbfm_gpu_get_turbo_history(self.h, first, n, _dp(out)))
        return out

    def turbo(self) -> np.ndarray:
        out = np.zeros(10)
        self._check(lib().tbfm_gpu_get_turbo(self.h, _dp(out)))
        return out

    def geometry(self, which: int) -> np.ndarray:
        ni, nj, nk = self.ni, self.nj, self.nk
        shapes = {0: (ni + 1, nj, nk, 3), 1: (ni, nj + 1, nk, 3), 2: (ni, nj, nk + 1, 3),
                  3: (ni, nj, nk), 4: (ni, nj, nk, 3)}
        out = np.zeros(shapes[which])
        self._check(lib().tbfm_gpu_get_geometry(self.h, which, _dp(out)))
        return out

    # -- profiling ---------------------------------------------------------------
    def profile(self, enable: bool = True, reset: bool = True) -> None:
        self._check(lib().tbfm_gpu_profile(self.h, int(enable), int(reset)))

    def profile_get(self):
        ms, n, b = C.c_double(), C.c_long(), C.c_double()
        self._check(lib().tbfm_gpu_profile_get(self.h, C.byref(ms), C.byref(n), C.byref(b)))
        return ms.value, n.value, b.value

    def profile_get_prestage(self):
        """Pre-stage kernel milliseconds over the profiled stages and the body-force data
        bytes one launch reads."""
        ms, b = C.c_double(), C.c_double()
        self._check(lib().tbfm_gpu_profile_get_prestage(self.h, C.byref(ms), C.byref(b)))
        return ms.value, b.value

    def tuning(self) -> str:
        return lib().tbfm_gpu_tuning(self.h).decode()

    def sync(self) -> None:
        self._check(lib().tbfm_gpu_sync(self.h))

    # -- batching / serving support ----------------------------------------------
    def device_bytes(self) -> int:
        b = C.c_long()
        self._check(lib().tbfm_gpu_case_device_bytes(self.h, C.byref(b)))
        return b.value

    def set_convergence(self, density_only: bool, drop: float) -> None:
        self._check(lib().tbfm_gpu_case_set_convergence(self.h, int(density_only), float(drop)))

    def status(self):
        """(iterations executed, convergence iteration or 0, stopped, last log residuals);
        raises SolverError if the case diverged."""
        it, conv, st = C.c_long(), C.c_long(), C.c_int()
        lr = np.zeros(5)
        rc = lib().tbfm_gpu_case_status(self.h, C.byref(it), C.byref(conv), C.byref(st), _dp(lr))
        self.iterations = it.value
        self.converged_at = conv.value
        self._check(rc)
        return it.value, conv.value, bool(st.value), lr


class Batch:
    """Cases advancing together: one launch per kernel over all members, replayed as one
    captured HIP graph per iteration (``tbfm_gpu_batch_*``).  Members keep their own
    state, counters and convergence test, and compute the same numbers alone and in any
    mix."""

    def __init__(self):
        h = C.c_void_p()
        rc = lib().tbfm_gpu_batch_create(C.byref(h))
        if rc != TBFM_OK:
            raise SolverError(lib().tbfm_gpu_batch_last_error(None).decode())
        self.h = h
        self.members = []

    def _check(self, rc):
        if rc != TBFM_OK:
            raise SolverError(lib().tbfm_gpu_batch_last_error(self.h).decode())

    def admit(self, s: Solver) -> None:
        self._check(lib().tbfm_gpu_batch_admit(self.h, s.h))
        self.members.append(s)

    def retire(self, s: Solver) -> None:
        self._check(lib().tbfm_gpu_batch_retire(self.h, s.h))
        self.members.remove(s)

    def __len__(self) -> int:
        return lib().tbfm_gpu_batch_size(self.h)

    def advance(self, n: int) -> None:
        self._check(lib().tbfm_gpu_batch_advance(self.h, n))

    def close(self) -> None:
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.tbfm_gpu_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        self.close()

