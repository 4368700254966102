"""Case description in the reference's JSON schema, plus seeded workload generators.

The schema is the one read by HyperFlow's ``JSONParser``
(src/fvm_2d_cart/simulation/parser/parser.cpp:13-488): ``model``, ``mesh.blocks[]``
with ``extents``/``x_cells``/``y_cells``/``boundaries``, ``initial_condition``,
``scheme`` and ``simulation``.  Keys the parser ignores are kept untouched.

``Case.initial_fields()`` fills the padded blocks exactly as
``MeshBlock::initialise_field_values`` does (mesh_block.cpp:215-257): interior
cells get the initial condition at the cell centroid, ghosts stay zero.  All
arithmetic is IEEE binary64 in the reference's evaluation order, so the arrays
are bit-identical to the reference's.
"""
from __future__ import annotations

import copy
import json
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

SIDES = ("bottom", "right", "top", "left")      # Mesh::apply_bcs order (mesh.cpp:33-38)
BC_KINDS = {"Transmissive": 0, "Reflective": 1, "Dirichlet": 2, "Internal": 3}
SCHEMES = {"Godunov": 0, "MUSCLHancock": 1}
LIMITERS = {"MinBee": 0, "VanLeer": 1}


def prim_to_cons(gamma: float, w) -> List[float]:
    """EulerIdealGasModel::prim_to_cons (euler_ideal_gas.cpp:96-109)."""
    rho, u, v, p = (float(x) for x in w)
    return [rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)]


@dataclass
class Block:
    x_left: float
    x_right: float
    y_bottom: float
    y_top: float
    nx: int
    ny: int
    boundaries: Dict[str, dict] = field(default_factory=dict)

    def to_json(self) -> dict:
        return {
            "extents": {"x_left": self.x_left, "x_right": self.x_right,
                        "y_bottom": self.y_bottom, "y_top": self.y_top},
            "boundaries": copy.deepcopy(self.boundaries),
            "x_cells": self.nx,
            "y_cells": self.ny,
        }

    @property
    def dx(self) -> float:
        return (self.x_right - self.x_left) / self.nx          # mesh_block.cpp:227-230

    @property
    def dy(self) -> float:
        return (self.y_top - self.y_bottom) / self.ny          # mesh_block.cpp:233-236


@dataclass
class Case:
    blocks: List[Block]
    initial_condition: dict
    scheme: str = "MUSCLHancock"
    limiter: Optional[str] = "MinBee"
    gamma: float = 1.4
    cfl: float = 0.5
    t_start: float = 0.0
    t_end: float = 1.0
    anim_duration: float = 1.0
    anim_fps: float = 30.0
    domain: Optional[dict] = None

    # ------------------------------------------------------------- schema --
    @property
    def ghost(self) -> int:
        return SCHEMES[self.scheme] + 1                        # parser.cpp:275-288

    def to_json(self) -> dict:
        sch = {"type": self.scheme, "riemann": "HLLCEuler", "ode": "ForwardEuler"}
        if self.scheme == "MUSCLHancock":
            sch["limiter"] = self.limiter
        dom = self.domain or {
            "x_left": min(b.x_left for b in self.blocks), "x_right": max(b.x_right for b in self.blocks),
            "y_bottom": min(b.y_bottom for b in self.blocks), "y_top": max(b.y_top for b in self.blocks)}
        return {
            "model": {"model": "EulerIdealGas", "gamma": self.gamma},
            "mesh": {"blocks": [b.to_json() for b in self.blocks]},
            "domain": dom,
            "initial_condition": copy.deepcopy(self.initial_condition),
            "scheme": sch,
            "simulation": {"cfl": self.cfl, "t_start": self.t_start, "t_end": self.t_end,
                           "animation": {"duration": self.anim_duration, "fps": self.anim_fps}},
        }

    def write_json(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump(self.to_json(), f, indent=1)

    @classmethod
    def from_json(cls, cfg: dict) -> "Case":
        blocks = []
        for bj in cfg["mesh"]["blocks"]:
            e = bj["extents"]
            blocks.append(Block(float(e["x_left"]), float(e["x_right"]), float(e["y_bottom"]), float(e["y_top"]),
                                int(bj["x_cells"]), int(bj["y_cells"]), copy.deepcopy(bj["boundaries"])))
        sch = cfg["scheme"]
        sim = cfg["simulation"]
        return cls(blocks=blocks, initial_condition=copy.deepcopy(cfg["initial_condition"]),
                   scheme=sch["type"], limiter=sch.get("limiter"), gamma=float(cfg["model"]["gamma"]),
                   cfl=float(sim["cfl"]), t_start=float(sim["t_start"]), t_end=float(sim["t_end"]),
                   anim_duration=float(sim["animation"]["duration"]), anim_fps=float(sim["animation"]["fps"]),
                   domain=copy.deepcopy(cfg.get("domain")))

    @classmethod
    def load(cls, path: str) -> "Case":
        with open(path) as f:
            return cls.from_json(json.load(f))

    # ------------------------------------------------ boundary conditions --
    def bc_tables(self):
        """Per block: kind[4], neighbour[4], conserved state[4][4] in side order
        bottom, right, top, left (parser.cpp:432-488)."""
        out = []
        for b in self.blocks:
            kinds, nbrs, states = [], [], []
            for s in SIDES:
                bc = b.boundaries[s]
                kind = BC_KINDS[bc["type"]]
                st = [0.0, 0.0, 0.0, 0.0]
                nb = -1
                if kind == 2:
                    fs = [float(x) for x in bc["flow_state"]]
                    st = prim_to_cons(self.gamma, fs) if bc.get("variables") == "Primitive" else fs
                elif kind == 3:
                    nb = int(bc["block"])
                kinds.append(kind)
                nbrs.append(nb)
                states.append(st)
            out.append((kinds, nbrs, states))
        return out

    def oracle_desc(self):
        """Flat descriptor arrays in the layout of oracle/hf_oracle.h."""
        idesc, ddesc = [], []
        for b, (kinds, nbrs, states) in zip(self.blocks, self.bc_tables()):
            idesc += [b.nx, b.ny] + kinds + nbrs
            ddesc += [b.x_left, b.x_right, b.y_bottom, b.y_top] + [x for st in states for x in st]
        return np.asarray(idesc, np.int32), np.asarray(ddesc, np.float64)

    # -------------------------------------------------- initial condition --
    def _ic_states(self):
        ic = self.initial_condition
        t = ic["type"]
        g = self.gamma
        prim = ic.get("variables") == "Primitive"

        def conv(key, force=False):
            st = [float(x) for x in ic[key]]
            return prim_to_cons(g, st) if (prim or force) else st

        if t == "Constant":                                    # parser.cpp:311-322
            return t, {"state": conv("init_state", force=True)}
        if t == "DoubleRiemannProblem":                        # parser.cpp:325-356
            return t, {"l": conv("init_left_state"), "m": conv("init_middle_state"),
                       "r": conv("init_right_state"),
                       "xlm": float(ic["x_lm_interface"]), "xmr": float(ic["x_mr_interface"])}
        if t == "QuadRiemannProblem":                          # parser.cpp:359-399
            return t, {"nw": conv("init_nw_state"), "ne": conv("init_ne_state"),
                       "sw": conv("init_sw_state"), "se": conv("init_se_state"),
                       "xi": float(ic["x_interface"]), "yi": float(ic["y_interface"])}
        if t == "SphericalRiemannProblem":                     # parser.cpp:402-428
            return t, {"in": conv("init_sphere_state", force=True), "ext": conv("init_ext_state", force=True),
                       "r": float(ic["radius"]), "x0": float(ic["x_origin"]), "y0": float(ic["y_origin"])}
        if t == "Fields":
            return t, {}
        raise ValueError(f"Initial condition '{t}' not supported")

    def initial_fields(self) -> List[np.ndarray]:
        """Padded conserved fields per block, shape (4, ny+2g, nx+2g)."""
        kind, st = self._ic_states()
        g = self.ghost
        out = []
        for b in self.blocks:
            tx, ty = b.nx + 2 * g, b.ny + 2 * g
            U = np.zeros((4, ty, tx), np.float64)
            dx, dy = b.dx, b.dy
            # MeshBlock::calculate_centroids (mesh_block.cpp:239-257)
            x0 = b.x_left - (float(g) - 0.5) * dx
            y0 = b.y_bottom - (float(g) - 0.5) * dy
            xs = x0 + np.arange(g, g + b.nx, dtype=np.float64) * dx
            ys = y0 + np.arange(g, g + b.ny, dtype=np.float64) * dy
            X, Y = np.meshgrid(xs, ys)                          # (ny, nx)
            if kind == "Constant":                              # constant.cpp:34-40
                vals = [np.full(X.shape, s) for s in st["state"]]
            elif kind == "DoubleRiemannProblem":                # doublerp.cpp:42-55
                left = X <= st["xlm"]
                mid = (X > st["xlm"]) & (X <= st["xmr"])
                vals = [np.where(left, st["l"][k], np.where(mid, st["m"][k], st["r"][k])) for k in range(4)]
            elif kind == "QuadRiemannProblem":                  # quadrp.cpp:49-64
                west = X <= st["xi"]
                south = Y <= st["yi"]
                vals = [np.where(west & south, st["sw"][k],
                                 np.where(west & ~south, st["nw"][k],
                                          np.where(~west & south, st["se"][k], st["ne"][k]))) for k in range(4)]
            elif kind == "SphericalRiemannProblem":             # sphericalrp.cpp:62-72
                inside = ((X - st["x0"]) * (X - st["x0"]) + (Y - st["y0"]) * (Y - st["y0"])) <= (st["r"] * st["r"])
                vals = [np.where(inside, st["in"][k], st["ext"][k]) for k in range(4)]
            else:
                raise ValueError("fields must be supplied explicitly for this case")
            for k in range(4):
                U[k, g:g + b.ny, g:g + b.nx] = vals[k]
            out.append(U)
        return out

    @property
    def interior_cells(self) -> int:
        return sum(b.nx * b.ny for b in self.blocks)


# ------------------------------------------------------------------ workloads --

def _bc(kind, **kw):
    d = {"type": kind}
    d.update(kw)
    return d


def quadrant_states(rng: np.random.Generator):
    """Seeded four-quadrant Riemann data: draw order SW, SE, NW, NE x (rho, u, v, p),
    rho,p ~ U[0.1, 2], u,v ~ U[-1, 1]."""
    states = {}
    for q in ("sw", "se", "nw", "ne"):
        rho = float(rng.uniform(0.1, 2.0))
        u = float(rng.uniform(-1.0, 1.0))
        v = float(rng.uniform(-1.0, 1.0))
        p = float(rng.uniform(0.1, 2.0))
        states[q] = [rho, u, v, p]
    return states


def quad_case(nx: int, ny: int, seed: int, scheme: str = "Godunov", limiter: Optional[str] = None,
              cfl: float = 0.8, bc: str = "Transmissive", gamma: float = 1.4,
              extents=(0.0, 1.0, 0.0, 1.0)) -> Case:
    """Single block with seeded quadrant data at the domain centre."""
    st = quadrant_states(np.random.default_rng(seed))
    xl, xr, yb, yt = extents
    blk = Block(xl, xr, yb, yt, nx, ny, {s: _bc(bc) for s in SIDES})
    ic = {"type": "QuadRiemannProblem", "x_interface": 0.5 * (xl + xr), "y_interface": 0.5 * (yb + yt),
          "variables": "Primitive", "init_nw_state": st["nw"], "init_ne_state": st["ne"],
          "init_sw_state": st["sw"], "init_se_state": st["se"]}
    return Case(blocks=[blk], initial_condition=ic, scheme=scheme, limiter=limiter, gamma=gamma, cfl=cfl)


def godunov_quadrants(seed: int = 0, n: int = 400) -> Case:
    """Godunov + HLLC, one n x n block on [0,1]^2, transmissive, CFL 0.8."""
    return quad_case(n, n, seed, scheme="Godunov", cfl=0.8)


def step_channel(n: int = 1024, limiter: str = "MinBee", cfl: float = 0.5) -> Case:
    """Eight n x n blocks (unit squares) tiling a channel with a solid step.

    Top row y in [1,2], x in [0,5] (blocks 3..7), bottom row y in [0,1], x in
    [0,3] (blocks 0..2); the step [3,5]x[0,1] is not meshed.  Left faces:
    Dirichlet Mach 3 inflow (rho=1.4, u=3, v=0, p=1, so c=1); the right face at
    x=5 is transmissive; walls and step faces are reflective; shared faces are
    Internal.  Initial condition: the inflow state everywhere.
    """
    inflow = _bc("Dirichlet", variables="Primitive", flow_state=[1.4, 3.0, 0.0, 1.0])
    wall = _bc("Reflective")
    blocks: List[Block] = []
    # bottom row: blocks 0, 1, 2
    for k in range(3):
        bnd = {"bottom": wall,
               "top": _bc("Internal", block=3 + k),
               "left": inflow if k == 0 else _bc("Internal", block=k - 1),
               "right": _bc("Internal", block=k + 1) if k < 2 else wall}
        blocks.append(Block(float(k), float(k + 1), 0.0, 1.0, n, n, bnd))
    # top row: blocks 3..7
    for k in range(5):
        bnd = {"top": wall,
               "bottom": _bc("Internal", block=k) if k < 3 else wall,
               "left": inflow if k == 0 else _bc("Internal", block=3 + k - 1),
               "right": _bc("Internal", block=3 + k + 1) if k < 4 else _bc("Transmissive")}
        blocks.append(Block(float(k), float(k + 1), 1.0, 2.0, n, n, bnd))
    ic = {"type": "Constant", "variables": "Primitive", "init_state": [1.4, 3.0, 0.0, 1.0]}
    return Case(blocks=blocks, initial_condition=ic, scheme="MUSCLHancock", limiter=limiter, cfl=cfl,
                domain={"x_left": 0.0, "x_right": 5.0, "y_bottom": 0.0, "y_top": 2.0})


def muscl_step_channel(n: int = 1024) -> Case:
    """MUSCL-Hancock + MinBee + HLLC, 8 blocks of n x n, step channel, CFL 0.5."""
    return step_channel(n, "MinBee", 0.5)


def perturbed_fields(case: Case, seed: int) -> List[np.ndarray]:
    """Quadrant fields with a multiplicative sinusoidal density perturbation
    rho <- rho * (1 + 0.1 sin(2 pi (kx x + ky y) + phi)), kx, ky ~ U{1..8},
    phi ~ U[0, 2 pi).  The random draws follow the quadrant draws."""
    rng = np.random.default_rng(seed)
    quadrant_states(rng)                      # same draws as quad_case(seed)
    kx = int(rng.integers(1, 9))
    ky = int(rng.integers(1, 9))
    phi = float(rng.uniform(0.0, 2.0 * math.pi))
    fields = case.initial_fields()
    g = case.ghost
    for b, U in zip(case.blocks, fields):
        dx, dy = b.dx, b.dy
        xs = (b.x_left - (float(g) - 0.5) * dx) + np.arange(g, g + b.nx) * dx
        ys = (b.y_bottom - (float(g) - 0.5) * dy) + np.arange(g, g + b.ny) * dy
        X, Y = np.meshgrid(xs, ys)
        f = 1.0 + 0.1 * np.sin(2.0 * math.pi * (kx * X + ky * Y) + phi)
        inner = (slice(None), slice(g, g + b.ny), slice(g, g + b.nx))
        Ui = U[inner]
        rho = Ui[0]
        u = Ui[1] / rho
        v = Ui[2] / rho
        p = (case.gamma - 1.0) * (Ui[3] - 0.5 * rho * (u * u + v * v))
        rho2 = rho * f
        Ui[0] = rho2
        Ui[1] = rho2 * u
        Ui[2] = rho2 * v
        Ui[3] = p / (case.gamma - 1.0) + 0.5 * rho2 * (u * u + v * v)
    return fields


def nonphysical_fields(case: Case, seed: int, block: int = 0) -> List[np.ndarray]:
    """perturbed_fields with three interior cells of one block made
    non-physical: negative total energy (so negative pressure and a NaN sound
    speed), zero density under non-zero momentum, and a NaN y-momentum.  They
    exercise the NaN paths of the scheme: HLLC's zero flux for NaN wave
    speeds, the limiters' comparisons with NaN ratios, and TimeStep skipping
    NaN terms."""
    fields = perturbed_fields(case, seed)
    g = case.ghost
    b = case.blocks[block]
    U = fields[block]

    def cell(fx: float, fy: float):
        return g + int(fy * b.ny), g + int(fx * b.nx)

    y, x = cell(0.3, 0.25)
    U[3, y, x] = -0.5 * abs(U[3, y, x])
    y, x = cell(0.55, 0.6)
    U[0, y, x] = 0.0
    U[1, y, x] = 0.3
    y, x = cell(0.2, 0.8)
    U[2, y, x] = np.nan
    return fields


def muscl_quadrants(seed: int = 0, n: int = 8192, limiter: str = "VanLeer") -> Case:
    """MUSCL-Hancock + Van Leer, one n x n block, transmissive, CFL 0.5
    (fields come from perturbed_fields)."""
    c = quad_case(n, n, seed, scheme="MUSCLHancock", limiter=limiter, cfl=0.5)
    return c


def batched_sizes(count: int = 512, seed: int = 0, lo: int = 64, hi: int = 512):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))) for _ in range(count)]
