"""Unstructured quadrilateral-cell Godunov solver (HyperFlow's src/fvm_2d_uns)
on the GPU: mesh files, test meshes, case files and the ctypes binding of
include/hyperflow_uns.h.

Mesh file (.hmsh, read by src/fvm_2d_uns/mesh/mesh_structured/
mesh_structured.cpp:23-122): a header line, the cell count, one line per
cell with its four vertices "x1 y1 x2 y2 x3 y3 x4 y4" (counter-clockwise),
the line "EDGES", the edge count, then one boundary code per cell edge in
cell order (edge k runs from vertex k to vertex k+1): 0 interior, -1
reflective, -2 transmissive, -3 inlet (src/fvm_2d_uns/scheme/boundary/
boundary.h:12).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Optional, Sequence, Tuple

import numpy as np

from . import native

REFLECTIVE, TRANSMISSIVE, INLET = -1, -2, -3


# ------------------------------------------------------------------ meshes --
def write_hmsh(path: str, quads: np.ndarray, edge_types: np.ndarray) -> None:
    """Write cells [n][4][2] and codes [n][4] in the .hmsh text format, each
    coordinate as its shortest decimal that reads back exactly (the text the
    reference generator writes, generate_block_mesh.py:217-243)."""
    quads = np.asarray(quads, np.float64)
    edge_types = np.asarray(edge_types, np.int64)
    with open(path, "w") as f:
        f.write("CELLS\n")
        f.write(f"{len(quads)}\n")
        for q in quads:
            f.write(" ".join(repr(float(v)) for v in q.reshape(-1)) + "\n")
        f.write("EDGES\n")
        f.write(f"{edge_types.size}\n")
        for t in edge_types.reshape(-1):
            f.write(f"{int(t)}\n")


def read_hmsh(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """Cells [n][4][2] and edge codes [n][4] of a .hmsh file."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[1], 0)
    k = 2
    cells = []
    while k < len(lines) and lines[k].strip() != "EDGES":
        vals = [float(v) for v in lines[k].split()]
        cells.append(vals[:8])
        k += 1
    m = int(lines[k + 1], 0)
    codes = [int(v, 0) for v in lines[k + 2:k + 2 + m]]
    quads = np.array(cells, np.float64).reshape(-1, 4, 2)
    if len(quads) != n:
        raise ValueError(f"{path}: header says {n} cells, found {len(quads)}")
    return quads, np.array(codes, np.int32).reshape(-1, 4)


def fma(a, b, c) -> np.ndarray:
    """Correctly rounded a*b + c on float64 arrays, from exact products and
    sums (Dekker's product, Knuth's TwoSum) with the low part rounded to odd
    (Boldo & Melquiond, "Emulation of FMA and correctly rounded sums: proved
    algorithms using rounding to odd", IEEE Trans. Computers 57(4), 2008).
    Valid away from overflow and underflow."""
    a, b, c = (np.asarray(v, np.float64) for v in (a, b, c))
    split = 134217729.0  # 2^27 + 1
    ah = a * split
    ah = ah - (ah - a)
    al = a - ah
    bh = b * split
    bh = bh - (bh - b)
    bl = b - bh
    ph = a * b
    pl = ((ah * bh - ph) + ah * bl + al * bh) + al * bl   # a*b = ph + pl exactly
    sh = c + ph
    z = sh - c
    sl = (c - (sh - z)) + (ph - z)                         # c + ph = sh + sl exactly
    v = sl + pl
    z = v - sl
    r = (sl - (v - z)) + (pl - z)                          # sl + pl = v + r exactly
    even = (v.view(np.int64) & 1) == 0
    v = np.where((r != 0) & even, np.nextafter(v, v + r), v)  # round to odd
    return sh + v


def block_mesh(blocks: Sequence[dict]) -> Tuple[np.ndarray, np.ndarray]:
    """Cells [n][4][2] and edge codes [n][4] of a block-structured mesh
    specification: a list of blocks {"coordinates": four corners (bottom
    left, bottom right, top right, top left), "cells": [n_eps, n_nu],
    "boundary_conditions": [south, east, north, west]}, as in the shipped
    scramjet_kumar_*_mesh.json files.

    BlockStructuredMeshGenerator (src/mesh_generation/generate_block_mesh.py:
    52-215): each block's corners span a bilinear map of the square
    [-1, 1]^2 with shape functions N_k = 0.25 (1 -+ eps)(1 -+ nu); cell
    (i, j) has block coordinates eps = i (2/n_eps) - 1, nu = j (2/n_nu) - 1,
    its vertices are the images of (eps, nu), (eps + d, nu), (eps + d,
    nu + d'), (eps, nu + d'); cells run block by block, eps-major; a cell's
    edge code is the block side's code where the edge lies on that side, else
    0.  The reference takes each coordinate as numpy's dot product of the
    four shape functions with the corner coordinates, which its BLAS
    evaluates as a chain of fused multiply-adds from 0.0 in corner order;
    the same chain here gives .hmsh files equal byte for byte."""
    quads, codes = [], []
    for b in blocks:
        ne, nn = int(b["cells"][0]), int(b["cells"][1])
        bx = [float(c[0]) for c in b["coordinates"]]
        by = [float(c[1]) for c in b["coordinates"]]
        bc = [int(v) for v in b["boundary_conditions"]]
        de, dn = 2.0 / ne, 2.0 / nn
        i, j = np.meshgrid(np.arange(ne, dtype=np.float64), np.arange(nn, dtype=np.float64), indexing="ij")
        eps = i.reshape(-1) * de - 1.0
        nu = j.reshape(-1) * dn - 1.0

        def image(e, u):
            N = (0.25 * (1.0 - e) * (1.0 - u), 0.25 * (1.0 + e) * (1.0 - u),
                 0.25 * (1.0 + e) * (1.0 + u), 0.25 * (1.0 - e) * (1.0 + u))
            x = np.zeros_like(e)
            y = np.zeros_like(e)
            for k in range(4):
                x = fma(N[k], bx[k], x)
                y = fma(N[k], by[k], y)
            return np.stack([x, y], axis=-1)

        q = np.stack([image(eps, nu), image(eps + de, nu), image(eps + de, nu + dn), image(eps, nu + dn)], axis=1)
        ii, jj = i.reshape(-1).astype(np.int64), j.reshape(-1).astype(np.int64)
        c = np.zeros((ne * nn, 4), np.int32)
        c[jj == 0, 0] = bc[0]
        c[ii == ne - 1, 1] = bc[1]
        c[jj == nn - 1, 2] = bc[2]
        c[ii == 0, 3] = bc[3]
        quads.append(q)
        codes.append(c)
    return np.concatenate(quads), np.concatenate(codes)


def ramp_channel(nx: int, ny: int, length: float = 3.0, height: float = 1.0, ramp_start: float = 1.0,
                 ramp_end: float = 2.0, ramp_height: float = 0.25, bottom: int = REFLECTIVE, top: int = REFLECTIVE,
                 left: int = INLET, right: int = TRANSMISSIVE):
    """A channel whose floor rises linearly between ramp_start and ramp_end
    (a compression ramp), meshed with nx x ny quadrilaterals: vertical cell
    columns, rows spaced evenly between the floor and the ceiling."""
    xs = np.linspace(0.0, length, nx + 1)
    floor = np.interp(xs, [0.0, ramp_start, ramp_end, length], [0.0, 0.0, ramp_height, ramp_height])
    X = np.repeat(xs[None, :], ny + 1, axis=0)
    t = np.linspace(0.0, 1.0, ny + 1)[:, None]
    Y = floor[None, :] + t * (height - floor[None, :])
    quads = np.empty((ny, nx, 4, 2))
    for j in range(ny):
        for i in range(nx):
            quads[j, i] = [[X[j, i], Y[j, i]], [X[j, i + 1], Y[j, i + 1]],
                           [X[j + 1, i + 1], Y[j + 1, i + 1]], [X[j + 1, i], Y[j + 1, i]]]
    codes = np.zeros((ny, nx, 4), np.int32)
    codes[0, :, 0] = bottom
    codes[:, -1, 1] = right
    codes[-1, :, 2] = top
    codes[:, 0, 3] = left
    return quads.reshape(-1, 4, 2), codes.reshape(-1, 4)


def box(nx: int, ny: int, size: float = 1.0, code: int = REFLECTIVE, skew: float = 0.0):
    """A square of nx x ny quadrilaterals with one boundary code on every
    side; skew > 0 shears the interior vertices (non-orthogonal cells)."""
    xs = np.linspace(0.0, size, nx + 1)
    ys = np.linspace(0.0, size, ny + 1)
    X, Y = np.meshgrid(xs, ys)
    if skew:
        inner = (slice(1, -1), slice(1, -1))
        X[inner] += skew * size / nx * np.sin(2 * np.pi * Y[inner] / size)
        Y[inner] += skew * size / ny * np.sin(2 * np.pi * X[inner] / size)
    quads = np.empty((ny, nx, 4, 2))
    for j in range(ny):
        for i in range(nx):
            quads[j, i] = [[X[j, i], Y[j, i]], [X[j, i + 1], Y[j, i + 1]],
                           [X[j + 1, i + 1], Y[j + 1, i + 1]], [X[j + 1, i], Y[j + 1, i]]]
    codes = np.zeros((ny, nx, 4), np.int32)
    codes[0, :, 0] = code
    codes[:, -1, 1] = code
    codes[-1, :, 2] = code
    codes[:, 0, 3] = code
    return quads.reshape(-1, 4, 2), codes.reshape(-1, 4)


def case_json(mesh_file: str, extents, inlet_prim, initial_condition: dict, cfl: float = 0.5, gamma: float = 1.4,
              t_end: float = 1.0) -> dict:
    """A case in the reference's unstructured JSON format
    (src/fvm_2d_uns/simulation/parser/parser.cpp:17-292)."""
    xl, xr, yb, yt = extents
    return {
        "model": {"model": "EulerIdealGas", "gamma": gamma},
        "mesh": {"type": "hyperflow_structured", "file": mesh_file,
                 "extents": {"x_left": xl, "x_right": xr, "y_bottom": yb, "y_top": yt},
                 "boundaries": {"inlet": list(inlet_prim)}},
        "domain": {"x_left": xl, "x_right": xr, "y_bottom": yb, "y_top": yt},
        "initial_condition": initial_condition,
        "scheme": {"type": "Godunov", "riemann": "HLLCEuler", "ode": "ForwardEuler"},
        "simulation": {"cfl": cfl, "t_start": 0.0, "t_end": t_end, "animation": {"duration": 1.0, "fps": 30.0}},
    }


def prim_to_cons(W, gamma: float) -> np.ndarray:
    """EulerIdealGasModel::prim_to_cons (model/euler_ideal_gas/euler_ideal_gas.cpp:96-109)."""
    rho, u, v, p = (float(w) for w in W)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.array([rho, rho * u, rho * v, E])


def initial_fields(quads: np.ndarray, ic: dict, gamma: float) -> np.ndarray:
    """Conserved fields [n][4] of an initial condition at the cell centroids
    (Mesh::initialise_field_values, mesh/mesh/mesh.cpp:60-68); the
    reference converts the states from primitive variables
    (parser.cpp:231-291)."""
    cen = centroids(quads)
    n = len(quads)
    if ic["type"] == "Constant":
        s = prim_to_cons(ic["init_state"], gamma)
        return np.repeat(s[None, :], n, axis=0)
    if ic["type"] == "SphericalRiemannProblem":
        r, xo, yo = float(ic["radius"]), float(ic["x_origin"]), float(ic["y_origin"])
        si = prim_to_cons(ic["init_sphere_state"], gamma)
        se = prim_to_cons(ic["init_ext_state"], gamma)
        inside = ((cen[:, 0] - xo) * (cen[:, 0] - xo) + (cen[:, 1] - yo) * (cen[:, 1] - yo)) <= (r * r)
        return np.where(inside[:, None], si[None, :], se[None, :])
    raise ValueError(f"initial condition {ic['type']!r} not supported")


def centroids(quads: np.ndarray) -> np.ndarray:
    """Polygon::calculate_centroid (geometry/polygon/polygon.cpp:179-210), in
    the reference's evaluation order (element-wise over all cells)."""
    q = np.asarray(quads, np.float64)
    x, y = q[:, :, 0], q[:, :, 1]
    vol = np.zeros(len(q))
    for i in range(4):
        j = (i + 1) % 4
        vol = vol + (x[:, i] * y[:, j] - x[:, j] * y[:, i])
    vol = vol * 0.5
    cx = np.zeros(len(q))
    cy = np.zeros(len(q))
    for i in range(4):
        j = (i + 1) % 4
        second = x[:, i] * y[:, j] - x[:, j] * y[:, i]
        cx = cx + (x[:, i] + x[:, j]) * second
        cy = cy + (y[:, i] + y[:, j]) * second
    return np.stack([cx / (6.0 * vol), cy / (6.0 * vol)], axis=1)


# ----------------------------------------------------------------- binding --
class hf_uns_desc(C.Structure):
    _fields_ = [("ncells", C.c_int32), ("pad", C.c_int32), ("vertices", C.POINTER(C.c_double)),
                ("edge_types", C.POINTER(C.c_int32)), ("gamma", C.c_double), ("cfl", C.c_double),
                ("inlet", C.c_double * 4)]


EXPORTS_UNS = ("hf_uns_create", "hf_uns_destroy", "hf_uns_cells", "hf_uns_connectivity", "hf_uns_geometry",
               "hf_uns_upload", "hf_uns_download", "hf_uns_time_step", "hf_uns_advance", "hf_uns_run",
               "hf_uns_synchronize", "hf_uns_dt_history", "hf_uns_profile", "hf_uns_kernel_stats", "hf_uns_bytes_per_cell")

_bound = False


def _lib():
    global _bound
    L = native.lib()
    if not _bound:
        vp = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.hf_uns_create.argtypes = [C.POINTER(hf_uns_desc), C.POINTER(vp)]
        L.hf_uns_destroy.argtypes = [vp]
        L.hf_uns_destroy.restype = None
        L.hf_uns_cells.argtypes = [vp]
        L.hf_uns_connectivity.argtypes = [vp, ip]
        L.hf_uns_geometry.argtypes = [vp, dp, dp]
        L.hf_uns_upload.argtypes = [vp, dp, C.c_size_t]
        L.hf_uns_download.argtypes = [vp, dp, C.c_size_t]
        L.hf_uns_time_step.argtypes = [vp, C.POINTER(C.c_double)]
        L.hf_uns_advance.argtypes = [vp, C.c_double]
        L.hf_uns_run.argtypes = [vp, C.c_int, C.c_int]
        L.hf_uns_synchronize.argtypes = [vp]
        L.hf_uns_dt_history.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
        L.hf_uns_profile.argtypes = [vp, C.c_int]
        L.hf_uns_bytes_per_cell.argtypes = [vp]
        L.hf_uns_kernel_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        _bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise native.HyperFlowError(f"{what} failed ({rc}): {native.lib().hf_last_error().decode()}")


class UnsSolver:
    """The unstructured Godunov solver on the GPU (include/hyperflow_uns.h)."""

    def __init__(self, quads, edge_types, inlet_cons=(0.0, 0.0, 0.0, 0.0), gamma: float = 1.4, cfl: float = 0.5):
        L = _lib()
        self._v = np.ascontiguousarray(quads, np.float64).reshape(-1)
        self._t = np.ascontiguousarray(edge_types, np.int32).reshape(-1)
        d = hf_uns_desc()
        d.ncells = len(self._v) // 8
        d.vertices = self._v.ctypes.data_as(C.POINTER(C.c_double))
        d.edge_types = self._t.ctypes.data_as(C.POINTER(C.c_int32))
        d.gamma = gamma
        d.cfl = cfl
        for k in range(4):
            d.inlet[k] = float(inlet_cons[k])
        h = C.c_void_p()
        _check(L.hf_uns_create(C.byref(d), C.byref(h)), "hf_uns_create")
        self._h = h
        self.ncells = d.ncells

    def close(self):
        if getattr(self, "_h", None):
            _lib().hf_uns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def connectivity(self) -> np.ndarray:
        out = np.empty((self.ncells, 4), np.int32)
        _check(_lib().hf_uns_connectivity(self._h, out), "hf_uns_connectivity")
        return out

    def geometry(self):
        cen = np.empty((self.ncells, 2))
        vol = np.empty(self.ncells)
        _check(_lib().hf_uns_geometry(self._h, cen, vol), "hf_uns_geometry")
        return cen, vol

    def upload(self, U) -> None:
        """U: conserved fields [n][4]."""
        a = np.ascontiguousarray(np.asarray(U, np.float64).T)
        _check(_lib().hf_uns_upload(self._h, a, a.size), "hf_uns_upload")

    def download(self) -> np.ndarray:
        a = np.empty((4, self.ncells))
        _check(_lib().hf_uns_download(self._h, a, a.size), "hf_uns_download")
        return np.ascontiguousarray(a.T)

    def time_step(self) -> float:
        dt = C.c_double()
        _check(_lib().hf_uns_time_step(self._h, C.byref(dt)), "hf_uns_time_step")
        return dt.value

    def advance(self, dt: float) -> None:
        _check(_lib().hf_uns_advance(self._h, float(dt)), "hf_uns_advance")

    def run(self, nsteps: int, graph: bool = False) -> None:
        """nsteps device-resident steps; graph=True captures them in a HIP
        graph (cached per step count and starting state)."""
        _check(_lib().hf_uns_run(self._h, int(nsteps), 1 if graph else 0), "hf_uns_run")

    def synchronize(self) -> None:
        _check(_lib().hf_uns_synchronize(self._h), "hf_uns_synchronize")

    def profile(self, enable: bool) -> None:
        _check(_lib().hf_uns_profile(self._h, int(bool(enable))), "hf_uns_profile")

    def kernel_stats(self):
        """(launches, summed milliseconds) of step launches since the last call."""
        n, ms = C.c_int64(), C.c_double()
        _check(_lib().hf_uns_kernel_stats(self._h, C.byref(n), C.byref(ms)), "hf_uns_kernel_stats")
        return n.value, ms.value

    def bytes_per_cell(self) -> int:
        """Bytes a step reads and writes per cell (each array once)."""
        b = _lib().hf_uns_bytes_per_cell(self._h)
        _check(0 if b > 0 else b, "hf_uns_bytes_per_cell")
        return b

    def dt_history(self) -> np.ndarray:
        n = C.c_int()
        _check(_lib().hf_uns_dt_history(self._h, None, 0, C.byref(n)), "hf_uns_dt_history")
        out = np.zeros(max(1, n.value))
        _check(_lib().hf_uns_dt_history(self._h, out.ctypes.data, n.value, C.byref(n)), "hf_uns_dt_history")
        return out[:n.value]


# --------------------------------------------------------------------- CLI --
def main(argv: Optional[Sequence[str]] = None) -> int:
    """python -m hyperflow_amd.uns mesh --block-json-file F --output-file O
    writes the .hmsh mesh of a block specification, with the options of the
    reference's generator (src/mesh_generation/generate_block_mesh.py:
    263-268).  The solver itself runs as hyperflow_amd/bin/hyperflow_uns."""
    import argparse
    p = argparse.ArgumentParser(prog="python -m hyperflow_amd.uns")
    sub = p.add_subparsers(dest="cmd", required=True)
    m = sub.add_parser("mesh", help="block JSON -> .hmsh")
    m.add_argument("--block-json-file", required=True)
    m.add_argument("--output-file", required=True)
    a = p.parse_args(argv)
    with open(a.block_json_file) as f:
        quads, codes = block_mesh(json.load(f))
    write_hmsh(a.output_file, quads, codes)
    print(f"wrote {a.output_file}: {len(quads)} cells")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
