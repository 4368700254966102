"""ctypes binding of the C ABI in include/hyperflow.h (libhyperflow_amd.so).

The library is the product: every call here goes to the HIP kernels.  There is
no CPU fallback -- if the library is missing or no GPU is present the calls
fail loudly.

A process that also uses torch on the GPU must import torch before the first
call here: torch bundles its own HIP runtime with the same soname as ROCm's,
and the process runs on whichever is loaded first (torch does not initialise
on another one).
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
           "hf_calls", "hf_profile", "hf_kernel_stats", "hf_fixup_stats", "hf_conservation", "hf_step_noalloc", "hf_last_error",
           "hf_version")
# include/hyperflow_learned.h
EXPORTS_LEARNED = ("hf_net_create", "hf_net_destroy", "hf_net_set_stream", "hf_net_set_desc", "hf_net_param_count",
                   "hf_net_set_params", "hf_net_get_params", "hf_net_reset_optimizer", "hf_net_forward",
                   "hf_net_gradients", "hf_net_train_step", "hf_net_synchronize", "hf_correct", "hf_run_fixed",
                   "hf_state_to_device", "hf_state_from_device", "hf_restrict2", "hf_stream")


class hf_block_desc(C.Structure):
    _fields_ = [("x_left", C.c_double), ("x_right", C.c_double), ("y_bottom", C.c_double), ("y_top", C.c_double),
                ("nx", C.c_int32), ("ny", C.c_int32), ("bc", C.c_int32 * 4), ("neighbour", C.c_int32 * 4),
                ("state", (C.c_double * 4) * 4), ("problem", C.c_int32), ("reserved", C.c_int32)]


class hf_case_desc(C.Structure):
    _fields_ = [("scheme", C.c_int32), ("limiter", C.c_int32), ("precision", C.c_int32), ("dt_mode", C.c_int32),
                ("gamma", C.c_double), ("cfl", C.c_double), ("nblocks", C.c_int32), ("tile_rows", C.c_int32),
                ("tail_rows", C.c_int32), ("fix_rows", C.c_int32), ("blocks", C.POINTER(hf_block_desc))]


class hf_net_desc(C.Structure):
    _fields_ = [("in_mean", C.c_float * 4), ("in_scale", C.c_float * 4), ("out_scale", C.c_float * 4),
                ("gamma", C.c_float), ("shock_threshold", C.c_float), ("max_output", C.c_float),
                ("reserved", C.c_float), ("seed", C.c_uint64)]


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
        L.hf_fixup_stats.argtypes = [vp, C.POINTER(C.c_int64)]
        L.hf_conservation.argtypes = [vp, C.c_int, _dp]
        L.hf_step_noalloc.argtypes = [C.POINTER(hf_case_desc), vp, vp, C.POINTER(C.c_int64), vp, vp]
        L.hf_last_error.restype = C.c_char_p
        L.hf_last_error.argtypes = []
        L.hf_version.restype = C.c_char_p
        L.hf_version.argtypes = []
        _fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.hf_net_create.argtypes = [C.POINTER(hf_net_desc), C.POINTER(vp)]
        L.hf_net_destroy.argtypes = [vp]
        L.hf_net_destroy.restype = None
        L.hf_net_set_stream.argtypes = [vp, vp]
        L.hf_net_set_desc.argtypes = [vp, C.POINTER(hf_net_desc)]
        L.hf_net_param_count.argtypes = []
        L.hf_net_param_count.restype = C.c_int64
        L.hf_net_set_params.argtypes = [vp, _fp, C.c_size_t]
        L.hf_net_get_params.argtypes = [vp, _fp, C.c_size_t]
        L.hf_net_reset_optimizer.argtypes = [vp]
        L.hf_net_forward.argtypes = [vp, vp, C.c_int, C.c_int, vp]
        L.hf_net_gradients.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp]
        L.hf_net_train_step.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_float, vp]
        L.hf_net_synchronize.argtypes = [vp]
        L.hf_correct.argtypes = [vp, vp, C.POINTER(C.c_int64)]
        L.hf_run_fixed.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        L.hf_state_to_device.argtypes = [vp, C.c_int, vp]
        L.hf_state_from_device.argtypes = [vp, C.c_int, vp]
        L.hf_restrict2.argtypes = [vp, C.c_int, vp]
        L.hf_stream.argtypes = [vp]
        L.hf_stream.restype = vp
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


def case_desc(cases, precision: str = "fp64", dt_mode: str = "compat", tile_rows: int = 0, tail_rows: int = 0,
              fix_rows: int = 0):
    """hf_case_desc for one or more cases (case k = problem k).  Returns the
    descriptor and the block array it points into (keep both alive)."""
    if isinstance(cases, Case):
        cases = [cases]
    c0 = cases[0]
    for c in cases[1:]:
        if (c.scheme, c.limiter, c.gamma, c.cfl) != (c0.scheme, c0.limiter, c0.gamma, c0.cfl):
            raise ValueError("all problems of one solver share scheme, limiter, gamma and cfl")
    descs = block_descs(cases)
    d = hf_case_desc()
    d.scheme = SCHEMES[c0.scheme]
    d.limiter = LIMITERS.get(c0.limiter or "MinBee", 0)
    d.precision = HF_FP64 if precision == "fp64" else HF_FP32
    d.dt_mode = HF_DT_COMPAT if dt_mode == "compat" else HF_DT_UNIFORM
    d.gamma = c0.gamma
    d.cfl = c0.cfl
    d.nblocks = len(descs)
    d.tile_rows = tile_rows
    d.tail_rows = tail_rows
    d.fix_rows = fix_rows
    d.blocks = C.cast(descs, C.POINTER(hf_block_desc))
    return d, descs


def step_noalloc(cases, d_in, d_out, calls, d_dt=None, precision: str = "fp64", dt_mode: str = "compat",
                 stream: int = 0, tile_rows: int = 0) -> None:
    """One step through hf_step_noalloc: d_in / d_out are device buffers (a
    tensor or an address) of the blocks' padded fields back to back; calls is
    an int64 numpy array of per-problem TimeStep counters, advanced in place."""
    d, keep = case_desc(cases, precision, dt_mode, tile_rows)
    if not (isinstance(calls, np.ndarray) and calls.dtype == np.int64 and calls.flags.c_contiguous):
        raise ValueError("calls must be a contiguous int64 numpy array")
    rc = lib().hf_step_noalloc(C.byref(d), C.c_void_p(_addr(d_in)), C.c_void_p(_addr(d_out)),
                               calls.ctypes.data_as(C.POINTER(C.c_int64)),
                               C.c_void_p(_addr(d_dt) if d_dt is not None else None), C.c_void_p(stream or None))
    del keep
    _check(rc, "hf_step_noalloc")


def _addr(t) -> int:
    return t.data_ptr() if hasattr(t, "data_ptr") else int(t)


class Solver:
    """One device context: a mesh of blocks (one or more independent problems)
    plus the scheme, advanced by the gfx950 kernels."""

    def __init__(self, cases, precision: str = "fp64", dt_mode: str = "compat", tile_rows: int = 0,
                 tail_rows: int = 0, fix_rows: int = 0):
        if isinstance(cases, Case):
            cases = [cases]
        self.cases: List[Case] = list(cases)
        c0 = self.cases[0]
        for c in self.cases[1:]:
            if (c.scheme, c.limiter, c.gamma, c.cfl) != (c0.scheme, c0.limiter, c0.gamma, c0.cfl):
                raise ValueError("all problems of one solver share scheme, limiter, gamma and cfl")
        self.blocks = [b for c in self.cases for b in c.blocks]
        d, self._descs = case_desc(self.cases, precision, dt_mode, tile_rows, tail_rows, fix_rows)
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

    def profile(self, enable, every: int = 1) -> None:
        """Time the update launches of every `every`-th step (off if not enable)."""
        _check(lib().hf_profile(self._h, int(every) if enable else 0), "hf_profile")

    def fixup_stats(self) -> dict:
        """Fix-up tallies of the compat MUSCL dt since the last call."""
        out = (C.c_int64 * 4)()
        _check(lib().hf_fixup_stats(self._h, out), "hf_fixup_stats")
        return {"steps": out[0], "skipped": out[1], "guess": out[2], "check": out[3]}

    def kernel_stats(self):
        n = C.c_int64()
        ms = C.c_double()
        _check(lib().hf_kernel_stats(self._h, C.byref(n), C.byref(ms)), "hf_kernel_stats")
        return n.value, ms.value

    def conservation(self, problem: int = 0) -> np.ndarray:
        out = np.zeros(4, np.float64)
        _check(lib().hf_conservation(self._h, problem, out), "hf_conservation")
        return out

    # ------------------------------------------- fixed dt, device transfers --
    def run_fixed(self, nsteps: int, dt: float, graph: bool = False) -> None:
        """nsteps steps with a fixed dt (update and MUSCL predictor)."""
        _check(lib().hf_run_fixed(self._h, int(nsteps), float(dt), HF_RUN_GRAPH if graph else 0), "hf_run_fixed")

    def state_to_device(self, b: int, dst) -> None:
        """Interior of block b into dense FP32 planes [4][ny][nx] at dst (a device tensor)."""
        _check(lib().hf_state_to_device(self._h, b, _ptr(dst)), "hf_state_to_device")

    def state_from_device(self, b: int, src) -> None:
        _check(lib().hf_state_from_device(self._h, b, _ptr(src)), "hf_state_from_device")

    def restrict2(self, b: int, dst) -> None:
        """2x2 block average of block b's interior into dense FP32 planes [4][ny/2][nx/2]."""
        _check(lib().hf_restrict2(self._h, b, _ptr(dst)), "hf_restrict2")

    def correct(self, net: "Net", count: bool = False) -> Optional[int]:
        """Learned correction of the current state in place; with count, the
        number of cells left at the plain state (synchronises)."""
        if count:
            n = C.c_int64()
            _check(lib().hf_correct(self._h, net._h, C.byref(n)), "hf_correct")
            return n.value
        _check(lib().hf_correct(self._h, net._h, None), "hf_correct")
        return None

    @property
    def stream(self) -> int:
        return lib().hf_stream(self._h) or 0


def _ptr(t) -> int:
    """Device address of a tensor (anything with data_ptr()) or an int."""
    if hasattr(t, "data_ptr"):
        if hasattr(t, "is_contiguous") and not t.is_contiguous():
            raise ValueError("device buffer must be contiguous")
        if hasattr(t, "dtype") and str(t.dtype) != "torch.float32":
            raise ValueError("device buffer must be float32")
        return t.data_ptr()
    return int(t)


class Net:
    """The learned-correction CNN (include/hyperflow_learned.h)."""

    def __init__(self, in_mean=(0.0,) * 4, in_scale=(1.0,) * 4, out_scale=(1.0,) * 4, gamma: float = 1.4,
                 shock_threshold: float = 0.0, seed: int = 0, max_output: float = 0.0):
        self._desc = self._make_desc(in_mean, in_scale, out_scale, gamma, shock_threshold, seed, max_output)
        h = C.c_void_p()
        _check(lib().hf_net_create(C.byref(self._desc), C.byref(h)), "hf_net_create")
        self._h = h

    @staticmethod
    def _make_desc(in_mean, in_scale, out_scale, gamma, shock_threshold, seed=0, max_output=0.0):
        d = hf_net_desc()
        for r in range(4):
            d.in_mean[r] = float(in_mean[r])
            d.in_scale[r] = float(in_scale[r])
            d.out_scale[r] = float(out_scale[r])
        d.gamma = float(gamma)
        d.shock_threshold = float(shock_threshold)
        d.max_output = float(max_output)
        d.seed = int(seed)
        return d

    def set_desc(self, in_mean, in_scale, out_scale, gamma: float = 1.4, shock_threshold: float = 0.0,
                 max_output: float = 0.0) -> None:
        self._desc = self._make_desc(in_mean, in_scale, out_scale, gamma, shock_threshold, 0, max_output)
        _check(lib().hf_net_set_desc(self._h, C.byref(self._desc)), "hf_net_set_desc")

    @property
    def desc(self) -> dict:
        d = self._desc
        return {"in_mean": list(d.in_mean), "in_scale": list(d.in_scale), "out_scale": list(d.out_scale),
                "gamma": d.gamma, "shock_threshold": d.shock_threshold, "max_output": d.max_output}

    def close(self):
        if getattr(self, "_h", None):
            lib().hf_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count() -> int:
        return int(lib().hf_net_param_count())

    def set_stream(self, stream: int) -> None:
        _check(lib().hf_net_set_stream(self._h, C.c_void_p(stream or None)), "hf_net_set_stream")

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().hf_net_set_params(self._h, p, p.size), "hf_net_set_params")

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count(), np.float32)
        _check(lib().hf_net_get_params(self._h, out, out.size), "hf_net_get_params")
        return out

    def reset_optimizer(self) -> None:
        _check(lib().hf_net_reset_optimizer(self._h), "hf_net_reset_optimizer")

    def forward(self, u, y) -> None:
        """y = net(normalised u); u, y: FP32 device tensors [4][ny][nx]."""
        _check(lib().hf_net_forward(self._h, _ptr(u), u.shape[2], u.shape[1], _ptr(y)), "hf_net_forward")

    def gradients(self, u, target):
        g = np.empty(self.param_count(), np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().hf_net_gradients(self._h, _ptr(u), _ptr(target), u.shape[2], u.shape[1], g.ctypes.data,
                                      loss.ctypes.data), "hf_net_gradients")
        return float(loss[0]), g

    def train_step(self, u, target, lr: float, loss_out=None) -> None:
        _check(lib().hf_net_train_step(self._h, _ptr(u), _ptr(target), u.shape[2], u.shape[1], float(lr),
                                       _ptr(loss_out) if loss_out is not None else None), "hf_net_train_step")

    def synchronize(self) -> None:
        _check(lib().hf_net_synchronize(self._h), "hf_net_synchronize")


def version() -> str:
    return lib().hf_version().decode()
