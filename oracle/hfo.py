"""ctypes loader for the CPU oracle (oracle/libhf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhf_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
REF_HARNESS_OMP = os.path.join(HERE, "_ref", "ref_harness_omp")   # OpenMP pragmas on: timing only

_lib = None
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.hfo_create.restype = C.c_void_p
        L.hfo_create.argtypes = [C.c_int, _ip, _dp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.hfo_destroy.argtypes = [C.c_void_p]
        L.hfo_ghost.argtypes = [C.c_void_p]
        L.hfo_set_fields.argtypes = [C.c_void_p, C.c_int, _dp]
        L.hfo_get_fields.argtypes = [C.c_void_p, C.c_int, _dp]
        L.hfo_set_threads.argtypes = [C.c_int]
        L.hfo_apply_bcs.argtypes = [C.c_void_p]
        L.hfo_time_step.argtypes = [C.c_void_p]
        L.hfo_time_step.restype = C.c_double
        L.hfo_calls.argtypes = [C.c_void_p]
        L.hfo_calls.restype = C.c_uint
        L.hfo_scheme_residual.argtypes = [C.c_void_p, C.c_int, _dp]
        L.hfo_step.argtypes = [C.c_void_p]
        L.hfo_step.restype = C.c_double
        L.hfo_run.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.hfo_run.restype = C.c_int
        L.hfo_dt_log.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.hfo_dt_log.restype = C.c_int
        L.hfo_cons_to_prim.argtypes = [C.c_double, _dp, _dp]
        L.hfo_prim_to_cons.argtypes = [C.c_double, _dp, _dp]
        L.hfo_cons_flux.argtypes = [C.c_double, _dp, C.c_int, _dp]
        L.hfo_hllc.argtypes = [C.c_double, _dp, _dp, C.c_int, _dp]
        L.hfo_limiter.argtypes = [C.c_int, C.c_double]
        L.hfo_limiter.restype = C.c_double
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().hfo_set_threads(int(n))


class OracleMesh:
    """The oracle's view of a Case (hyperflow_amd.case.Case)."""

    def __init__(self, case, fields: Optional[Sequence[np.ndarray]] = None):
        from hyperflow_amd.case import LIMITERS, SCHEMES
        self.case = case
        idesc, ddesc = case.oracle_desc()
        lim = LIMITERS.get(case.limiter or "MinBee", 0)
        self._h = lib().hfo_create(len(case.blocks), idesc, ddesc, SCHEMES[case.scheme], lim,
                                   float(case.gamma), float(case.cfl))
        self.g = case.ghost
        if fields is None:
            fields = case.initial_fields()
        for b, U in enumerate(fields):
            self.set_fields(b, U)

    def close(self):
        if self._h:
            lib().hfo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, b):
        blk = self.case.blocks[b]
        return (4, blk.ny + 2 * self.g, blk.nx + 2 * self.g)

    def set_fields(self, b, U):
        U = np.ascontiguousarray(U, np.float64)
        assert U.shape == self.shape(b)
        lib().hfo_set_fields(self._h, b, U)

    def fields(self, b) -> np.ndarray:
        out = np.empty(self.shape(b), np.float64)
        lib().hfo_get_fields(self._h, b, out)
        return out

    def all_fields(self) -> List[np.ndarray]:
        return [self.fields(b) for b in range(len(self.case.blocks))]

    def apply_bcs(self):
        lib().hfo_apply_bcs(self._h)

    def time_step(self) -> float:
        return lib().hfo_time_step(self._h)

    @property
    def calls(self) -> int:
        return lib().hfo_calls(self._h)

    def residual(self, b) -> np.ndarray:
        out = np.empty(self.shape(b), np.float64)
        lib().hfo_scheme_residual(self._h, b, out)
        return out

    def step(self) -> float:
        return lib().hfo_step(self._h)

    def run(self, nsteps: int):
        dts = np.zeros(max(nsteps, 1), np.float64)
        cap = nsteps * (1 + len(self.case.blocks)) + 1
        calls = np.zeros(cap, np.float64)
        made = lib().hfo_run(self._h, nsteps, dts, calls, cap)
        return dts[:nsteps], calls[:made]

    def dt_log(self) -> np.ndarray:
        n = lib().hfo_dt_log(self._h, None, 0)
        out = np.zeros(max(n, 1), np.float64)
        lib().hfo_dt_log(self._h, out.ctypes.data, n)
        return out[:n]


# ----------------------------------------------------------- point-wise --

def cons_to_prim(gamma, U):
    out = np.zeros(4)
    lib().hfo_cons_to_prim(gamma, np.ascontiguousarray(U, np.float64), out)
    return out


def prim_to_cons(gamma, W):
    out = np.zeros(4)
    lib().hfo_prim_to_cons(gamma, np.ascontiguousarray(W, np.float64), out)
    return out


def cons_flux(gamma, U, direction):
    out = np.zeros(4)
    lib().hfo_cons_flux(gamma, np.ascontiguousarray(U, np.float64), int(direction), out)
    return out


def hllc(gamma, UL, UR, direction):
    out = np.zeros(4)
    lib().hfo_hllc(gamma, np.ascontiguousarray(UL, np.float64), np.ascontiguousarray(UR, np.float64),
                   int(direction), out)
    return out


def limiter(kind, r):
    return lib().hfo_limiter(int(kind), float(r))


# --------------------------------------------------- reference harness --

def have_reference() -> bool:
    return os.path.exists(REF_HARNESS)


def run_reference(case, nsteps: int, fields: Optional[Sequence[np.ndarray]] = None, timeout: float = 3600,
                  omp_threads: int = 0):
    """Run the reference's own classes (oracle/_ref/ref_harness) on a case.
    Returns (fields per block, all TimeStep calls, step dts, elapsed seconds).
    With omp_threads the build with the reference's OpenMP pragmas enabled
    runs on that many threads: for timing only (its TimeStep minimum is
    updated without a reduction)."""
    if not have_reference():
        raise RuntimeError("reference harness not built")
    harness, env = REF_HARNESS, None
    if omp_threads:
        if not os.path.exists(REF_HARNESS_OMP):
            raise RuntimeError("OpenMP reference harness not built")
        harness, env = REF_HARNESS_OMP, dict(os.environ, OMP_NUM_THREADS=str(omp_threads))
    with tempfile.TemporaryDirectory() as td:
        cj = os.path.join(td, "case.json")
        out = os.path.join(td, "out.bin")
        case.write_json(cj)
        args = [harness, cj, str(nsteps), out]
        if fields is not None:
            fin = os.path.join(td, "fields.bin")
            with open(fin, "wb") as f:
                for U in fields:
                    f.write(np.ascontiguousarray(U, np.float64).tobytes())
            args.append(fin)
        subprocess.run(args, check=True, stdout=subprocess.DEVNULL, timeout=timeout, env=env)
        return read_ref_output(out)


def read_ref_output(path):
    with open(path, "rb") as f:
        data = f.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from(fmt, data, off)
        off += struct.calcsize(fmt)
        return v

    (nb,) = take("<i")
    dims = [take("<iii") for _ in range(nb)]
    fields = []
    for nx, ny, g in dims:
        n = 4 * (nx + 2 * g) * (ny + 2 * g)
        a = np.frombuffer(data, np.float64, n, off).reshape(4, ny + 2 * g, nx + 2 * g).copy()
        off += 8 * n
        fields.append(a)
    (nc,) = take("<i")
    calls = np.frombuffer(data, np.float64, nc, off).copy()
    off += 8 * nc
    (ns,) = take("<i")
    steps = np.frombuffer(data, np.float64, ns, off).copy()
    off += 8 * ns
    (elapsed,) = take("<d")
    return fields, calls, steps, elapsed


# ------------------------------------------- unstructured solver (fvm_2d_uns) --
REF_UNS_HARNESS = os.path.join(HERE, "_ref", "ref_uns_harness")
_uns_bound = False


def _uns_lib():
    global _uns_bound
    L = lib()
    if not _uns_bound:
        L.hfo_uns_create.restype = C.c_void_p
        L.hfo_uns_create.argtypes = [C.c_int, _dp, _ip, C.c_double, C.c_double, _dp]
        L.hfo_uns_destroy.argtypes = [C.c_void_p]
        L.hfo_uns_connectivity.argtypes = [C.c_void_p, _ip]
        L.hfo_uns_geometry.argtypes = [C.c_void_p, _dp, _dp]
        L.hfo_uns_set_fields.argtypes = [C.c_void_p, _dp]
        L.hfo_uns_get_fields.argtypes = [C.c_void_p, _dp]
        L.hfo_uns_time_step.argtypes = [C.c_void_p]
        L.hfo_uns_time_step.restype = C.c_double
        L.hfo_uns_advance.argtypes = [C.c_void_p, C.c_double]
        _uns_bound = True
    return L


class UnsOracle:
    """The C restatement of the reference's unstructured Godunov solver."""

    def __init__(self, quads, codes, inlet_cons, gamma=1.4, cfl=0.5):
        L = _uns_lib()
        self._q = np.ascontiguousarray(quads, np.float64).reshape(-1)
        self._c = np.ascontiguousarray(codes, np.int32).reshape(-1)
        self.n = len(self._q) // 8
        self._h = L.hfo_uns_create(self.n, self._q, self._c, gamma, cfl,
                                   np.ascontiguousarray(inlet_cons, np.float64))

    def close(self):
        if getattr(self, "_h", None):
            _uns_lib().hfo_uns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def connectivity(self):
        out = np.empty((self.n, 4), np.int32)
        _uns_lib().hfo_uns_connectivity(self._h, out)
        return out

    def geometry(self):
        cen = np.empty((self.n, 2))
        vol = np.empty(self.n)
        _uns_lib().hfo_uns_geometry(self._h, cen, vol)
        return cen, vol

    def set_fields(self, U):
        _uns_lib().hfo_uns_set_fields(self._h, np.ascontiguousarray(U, np.float64))

    def fields(self):
        out = np.empty((self.n, 4))
        _uns_lib().hfo_uns_get_fields(self._h, out)
        return out

    def time_step(self):
        return _uns_lib().hfo_uns_time_step(self._h)

    def advance(self, dt):
        _uns_lib().hfo_uns_advance(self._h, dt)

    def run(self, nsteps):
        dts = []
        for _ in range(nsteps):
            dt = self.time_step()
            self.advance(dt)
            dts.append(dt)
        return np.array(dts)


def have_uns_reference() -> bool:
    return os.path.exists(REF_UNS_HARNESS)


def run_uns_reference(case_json: dict, nsteps: int, timeout: float = 3600, fields=None):
    """Run the reference's unstructured solver (oracle/_ref/ref_uns_harness)
    on a case dict whose mesh.file is an absolute .hmsh path, from the case's
    initial condition or from `fields` [n][4].  Returns a dict with
    connectivity [n][4], centroids, volumes, fields [n][4], dts, elapsed."""
    import json
    if not have_uns_reference():
        raise RuntimeError("reference unstructured harness not built")
    with tempfile.TemporaryDirectory() as td:
        cj = os.path.join(td, "case.json")
        out = os.path.join(td, "out.bin")
        with open(cj, "w") as f:
            json.dump(case_json, f)
        args = [REF_UNS_HARNESS, cj, str(nsteps), out]
        if fields is not None:
            args.append(os.path.join(td, "fields.bin"))
            np.ascontiguousarray(fields, np.float64).tofile(args[-1])
        subprocess.run(args, check=True, stdout=subprocess.DEVNULL, timeout=timeout)
        data = open(out, "rb").read()
    nc, ns = struct.unpack_from("<ii", data, 0)
    off = 8
    rec = np.frombuffer(data, np.int32, nc * 12, off).reshape(nc, 4, 3)
    off += nc * 48
    cen = np.frombuffer(data, np.float64, nc * 2, off).reshape(nc, 2).copy()
    off += nc * 16
    vol = np.frombuffer(data, np.float64, nc, off).copy()
    off += nc * 8
    U = np.frombuffer(data, np.float64, nc * 4, off).reshape(nc, 4).copy()
    off += nc * 32
    dts = np.frombuffer(data, np.float64, ns, off).copy()
    off += ns * 8
    (elapsed,) = struct.unpack_from("<d", data, off)
    return {"connectivity": rec[:, :, 1].copy(), "centroids": cen, "volumes": vol, "fields": U, "dts": dts,
            "elapsed": elapsed}
