"""ctypes binding of the C ABI in include/hyperflow.h (libhyperflow_amd.so).

The library is the product: every call here goes to the HIP kernels.  There is
no CPU fallback -- if the library is missing or no GPU is present the calls
fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from .case import BC_KINDS, LIMITERS, SCHEMES, Case

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhyperflow_amd.so")

HF_FP64, HF_FP32 = 0, 1
HF_DT_COMPAT, HF_DT_UNIFORM = 0, 1
HF_RUN_GRAPH = 1

EXPORTS = ("hf_create", "hf_destroy", "hf_ghost_cells", "hf_upload", "hf_download", "hf_apply_bcs",
           "hf_time_step", "hf_scheme_residual", "hf_advance", "hf_run", "hf_synchronize", "hf_dt_history",
           "hf_calls", "hf_profile", "hf_kernel_stats", "hf_conservation", "hf_last_error", "hf_version")


class hf_block_desc(C.Structure):
    _fields_ = [("x_left", C.c_double), ("x_right", C.c_double), ("y_bottom", C.c_double), ("y_top", C.c_double),
                ("nx", C.c_int32), ("ny", C.c_int32), ("bc", C.c_int32 * 4), ("neighbour", C.c_int32 * 4),
                ("state", (C.c_double * 4) * 4), ("problem", C.c_int32), ("reserved", C.c_int32)]


class hf_case_desc(C.Structure):
    _fields_ = [("scheme", C.c_int32), ("limiter", C.c_int32), ("precision", C.c_int32), ("dt_mode", C.c_int32),
                ("gamma", C.c_double), ("cfl", C.c_double), ("nblocks", C.c_int32), ("tile_rows", C.c_int32),
                ("blocks", C.POINTER(hf_block_desc))]


class HyperFlowError(RuntimeError):
    pass


_lib = None
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HyperFlowError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.hf_create.argtypes = [C.POINTER(hf_case_desc), C.POINTER(vp)]
        L.hf_destroy.argtypes = [vp]
        L.hf_destroy.restype = None
        L.hf_ghost_cells.argtypes = [vp]
        L.hf_upload.argtypes = [vp, C.c_int, _dp, C.c_size_t]
        L.hf_download.argtypes = [vp, C.c_int, _dp, C.c_size_t]
        L.hf_apply_bcs.argtypes = [vp]
        L.hf_time_step.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.hf_scheme_residual.argtypes = [vp, C.c_int, _dp, C.c_size_t]
        L.hf_advance.argtypes = [vp, _dp]
        L.hf_run.argtypes = [vp, C.c_int, C.c_int]
        L.hf_synchronize.argtypes = [vp]
        L.hf_dt_history.argtypes = [vp, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.hf_calls.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.hf_profile.argtypes = [vp, C.c_int]
        L.hf_kernel_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.hf_conservation.argtypes = [vp, C.c_int, _dp]
        L.hf_last_error.restype = C.c_char_p
        L.hf_last_error.argtypes = []
        L.hf_version.restype = C.c_char_p
        L.hf_version.argtypes = []
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().hf_last_error().decode()
        raise HyperFlowError(f"{what} failed ({rc}): {msg}")


def block_descs(cases: Sequence[Case]):
    """hf_block_desc array for one or more cases; case k becomes problem k."""
    descs = []
    base = 0
    for p, case in enumerate(cases):
        for b, (blk, (kinds, nbrs, states)) in enumerate(zip(case.blocks, case.bc_tables())):
            d = hf_block_desc()
            d.x_left, d.x_right, d.y_bottom, d.y_top = blk.x_left, blk.x_right, blk.y_bottom, blk.y_top
            d.nx, d.ny = blk.nx, blk.ny
            for s in range(4):
                d.bc[s] = kinds[s]
                d.neighbour[s] = nbrs[s] + base if nbrs[s] >= 0 else -1
                for v in range(4):
                    d.state[s][v] = states[s][v]
            d.problem = p
            descs.append(d)
        base += len(case.blocks)
    arr = (hf_block_desc * len(descs))(*descs)
    return arr


class Solver:
    """One device context: a mesh of blocks (one or more independent problems)
    plus the scheme, advanced by the gfx950 kernels."""

    def __init__(self, cases, precision: str = "fp64", dt_mode: str = "compat", tile_rows: int = 0):
        if isinstance(cases, Case):
            cases = [cases]
        self.cases: List[Case] = list(cases)
        c0 = self.cases[0]
        for c in self.cases[1:]:
            if (c.scheme, c.limiter, c.gamma, c.cfl) != (c0.scheme, c0.limiter, c0.gamma, c0.cfl):
                raise ValueError("all problems of one solver share scheme, limiter, gamma and cfl")
        self.blocks = [b for c in self.cases for b in c.blocks]
        self._descs = block_descs(self.cases)
        d = hf_case_desc()
        d.scheme = SCHEMES[c0.scheme]
        d.limiter = LIMITERS.get(c0.limiter or "MinBee", 0)
        d.precision = HF_FP64 if precision == "fp64" else HF_FP32
        d.dt_mode = HF_DT_COMPAT if dt_mode == "compat" else HF_DT_UNIFORM
        d.gamma = c0.gamma
        d.cfl = c0.cfl
        d.nblocks = len(self._descs)
        d.tile_rows = tile_rows
        d.blocks = C.cast(self._descs, C.POINTER(hf_block_desc))
        self._desc = d
        h = C.c_void_p()
        _check(lib().hf_create(C.byref(d), C.byref(h)), "hf_create")
        self._h = h
        self.g = lib().hf_ghost_cells(h)
        self.precision = precision

    def close(self):
        if getattr(self, "_h", None):
            lib().hf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ----------------------------------------------------------- fields --
    def shape(self, b: int):
        blk = self.blocks[b]
        return (4, blk.ny + 2 * self.g, blk.nx + 2 * self.g)

    def upload(self, b: int, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, np.float64)
        if U.shape != self.shape(b):
            raise ValueError(f"block {b}: expected shape {self.shape(b)}, got {U.shape}")
        _check(lib().hf_upload(self._h, b, U, U.size), "hf_upload")

    def upload_all(self, fields: Sequence[np.ndarray]) -> None:
        for b, U in enumerate(fields):
            self.upload(b, U)

    def download(self, b: int) -> np.ndarray:
        out = np.empty(self.shape(b), np.float64)
        _check(lib().hf_download(self._h, b, out, out.size), "hf_download")
        return out

    def download_all(self) -> List[np.ndarray]:
        return [self.download(b) for b in range(len(self.blocks))]

    # ---------------------------------------------------------- stepping --
    def apply_bcs(self) -> None:
        _check(lib().hf_apply_bcs(self._h), "hf_apply_bcs")

    def time_step(self, problem: int = 0) -> float:
        dt = C.c_double()
        _check(lib().hf_time_step(self._h, problem, C.byref(dt)), "hf_time_step")
        return dt.value

    def residual(self, b: int) -> np.ndarray:
        out = np.empty(self.shape(b), np.float64)
        _check(lib().hf_scheme_residual(self._h, b, out, out.size), "hf_scheme_residual")
        return out

    def advance(self, dts) -> None:
        dts = np.ascontiguousarray(np.atleast_1d(np.asarray(dts, np.float64)))
        _check(lib().hf_advance(self._h, dts), "hf_advance")

    def run(self, nsteps: int, graph: bool = False) -> None:
        _check(lib().hf_run(self._h, int(nsteps), HF_RUN_GRAPH if graph else 0), "hf_run")

    def synchronize(self) -> None:
        _check(lib().hf_synchronize(self._h), "hf_synchronize")

    def dt_history(self) -> np.ndarray:
        n = C.c_int()
        _check(lib().hf_dt_history(self._h, None, 0, C.byref(n)), "hf_dt_history")
        out = np.zeros(max(n.value, 1), np.float64)
        _check(lib().hf_dt_history(self._h, out.ctypes.data, n.value, C.byref(n)), "hf_dt_history")
        return out[:n.value]

    def calls(self, problem: int = 0) -> int:
        v = C.c_int64()
        _check(lib().hf_calls(self._h, problem, C.byref(v)), "hf_calls")
        return v.value

    def profile(self, enable: bool) -> None:
        _check(lib().hf_profile(self._h, int(enable)), "hf_profile")

    def kernel_stats(self):
        n = C.c_int64()
        ms = C.c_double()
        _check(lib().hf_kernel_stats(self._h, C.byref(n), C.byref(ms)), "hf_kernel_stats")
        return n.value, ms.value

    def conservation(self, problem: int = 0) -> np.ndarray:
        out = np.zeros(4, np.float64)
        _check(lib().hf_conservation(self._h, problem, out), "hf_conservation")
        return out


def version() -> str:
    return lib().hf_version().decode()
