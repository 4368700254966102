"""The oracle of the unstructured solver (hfo.UnsOracle, the C restatement of
src/fvm_2d_uns) and the host-side mesh tools of hyperflow_amd.uns, on the CPU.

Pins: the reference solver's own outputs (tests/golden/uns_*.npz, written by
oracle/_ref/ref_uns_harness), the reference mesh generator's .hmsh text for
the shipped scramjet mesh (uns_kumar_coarse.json), and what the mathematics
fixes: on an axis-aligned box the scheme is the Cartesian Godunov update with
HLLC fluxes, closed walls conserve mass, momentum-free uniform states and
uniform flow along walls are steady, the first TimeStep call is 0.2*cfl*dx /
(|u| + c), geometry has closed forms."""
import glob
import hashlib
import json
import math
import os

import numpy as np
import pytest

import hfo
from hyperflow_amd import uns as U

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "uns_*.npz")))
KUMAR = os.path.join(HERE, "golden", "uns_kumar_coarse.json")
GAMMA = 1.4


def golden(path):
    g = np.load(path)
    cj = json.loads(str(g["case_json"]))
    inlet = U.prim_to_cons(cj["mesh"]["boundaries"]["inlet"], cj["model"]["gamma"])
    return g, cj, inlet


def test_goldens_present():
    assert len(GOLDEN) >= 3


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_matches_reference(path):
    g, cj, inlet = golden(path)
    o = hfo.UnsOracle(g["quads"], g["codes"], inlet, cj["model"]["gamma"], cj["simulation"]["cfl"])
    np.testing.assert_array_equal(o.connectivity(), g["connectivity"])
    cen, vol = o.geometry()
    np.testing.assert_array_equal(cen, g["centroids"])
    np.testing.assert_array_equal(vol, g["volumes"])
    init = U.initial_fields(g["quads"], cj["initial_condition"], cj["model"]["gamma"])
    if "nonphysical" in os.path.basename(path):
        # three cells replaced after the initial condition (make_golden.py)
        same = (init == g["init"]) | (np.isnan(init) & np.isnan(g["init"]))
        assert int(np.sum(~np.all(same, axis=1))) == 3
        init = g["init"]
    np.testing.assert_array_equal(init, g["init"])
    o.set_fields(init)
    dts = o.run(int(g["steps"]))
    np.testing.assert_array_equal(dts, g["dts"])
    np.testing.assert_array_equal(o.fields(), g["fields"])


def cart_godunov_step(W, nx, ny, dx, dy, dt):
    """One unsplit Godunov step on an axis-aligned nx x ny grid with
    transmissive sides, from the oracle's point-wise HLLC: the textbook
    update U - dt/dx (F_{i+1/2} - F_{i-1/2}) - dt/dy (G_{j+1/2} - G_{j-1/2})."""
    Ug = W.reshape(ny, nx, 4)
    out = np.empty_like(Ug)
    for j in range(ny):
        for i in range(nx):
            c = Ug[j, i]
            lft = Ug[j, i - 1] if i > 0 else c
            rgt = Ug[j, i + 1] if i < nx - 1 else c
            bot = Ug[j - 1, i] if j > 0 else c
            top = Ug[j + 1, i] if j < ny - 1 else c
            fx = hfo.hllc(GAMMA, c, rgt, 0) - hfo.hllc(GAMMA, lft, c, 0)
            fy = hfo.hllc(GAMMA, c, top, 1) - hfo.hllc(GAMMA, bot, c, 1)
            out[j, i] = c - dt * fx / dx - dt * fy / dy
    return out.reshape(-1, 4)


def test_axis_aligned_box_is_cartesian_godunov():
    nx, ny = 9, 7
    q, c = U.box(nx, ny, size=1.0, code=U.TRANSMISSIVE)
    rng = np.random.default_rng(3)
    W = np.stack([rng.uniform(0.5, 2.0, nx * ny), rng.uniform(-0.5, 0.5, nx * ny),
                  rng.uniform(-0.5, 0.5, nx * ny), rng.uniform(0.5, 2.0, nx * ny)], axis=1)
    U0 = np.array([U.prim_to_cons(w, GAMMA) for w in W])
    o = hfo.UnsOracle(q, c, [0, 0, 0, 0], GAMMA, 0.5)
    o.set_fields(U0)
    dt = 1e-3
    o.advance(dt)
    want = cart_godunov_step(U0, nx, ny, 1.0 / nx, 1.0 / ny, dt)
    np.testing.assert_allclose(o.fields(), want, rtol=1e-12, atol=1e-13)


def test_reflective_box_conserves_mass_and_energy():
    q, c = U.box(14, 11, code=U.REFLECTIVE, skew=0.25)
    ic = {"type": "SphericalRiemannProblem", "radius": 0.25, "x_origin": 0.45, "y_origin": 0.5,
          "init_sphere_state": [1.0, 0.3, -0.2, 5.0], "init_ext_state": [0.3, 0.0, 0.0, 0.2]}
    U0 = U.initial_fields(q, ic, GAMMA)
    o = hfo.UnsOracle(q, c, [0, 0, 0, 0], GAMMA, 0.8)
    o.set_fields(U0)
    _, vol = o.geometry()
    before = (U0 * vol[:, None]).sum(axis=0)
    o.run(30)
    after = (o.fields() * vol[:, None]).sum(axis=0)
    assert abs(after[0] - before[0]) <= 1e-13 * before[0]
    assert abs(after[3] - before[3]) <= 1e-13 * before[3]


def test_uniform_states_are_steady():
    # a fluid at rest in a sheared, walled box, and a uniform stream along a
    # straight walled channel fed by its own state at the inlet
    q, c = U.box(10, 8, code=U.REFLECTIVE, skew=0.3)
    rest = U.prim_to_cons([1.2, 0.0, 0.0, 0.9], GAMMA)
    o = hfo.UnsOracle(q, c, [0, 0, 0, 0], GAMMA, 0.5)
    o.set_fields(np.repeat(rest[None], len(q), axis=0))
    o.run(10)
    np.testing.assert_allclose(o.fields(), np.repeat(rest[None], len(q), axis=0), rtol=1e-14, atol=1e-14)
    q, c = U.ramp_channel(12, 5, ramp_height=0.0)
    stream = U.prim_to_cons([1.4, 3.0, 0.0, 1.0], GAMMA)
    o = hfo.UnsOracle(q, c, stream, GAMMA, 0.5)
    o.set_fields(np.repeat(stream[None], len(q), axis=0))
    o.run(10)
    np.testing.assert_allclose(o.fields(), np.repeat(stream[None], len(q), axis=0), rtol=1e-14, atol=1e-13)


def test_first_time_step_closed_form():
    nx = 8
    q, c = U.box(nx, 4, size=1.0, code=U.TRANSMISSIVE)
    rho, u, v, p = 1.0, 0.6, -0.8, 1.0
    o = hfo.UnsOracle(q, c, [0, 0, 0, 0], GAMMA, 0.5)
    o.set_fields(np.repeat(U.prim_to_cons([rho, u, v, p], GAMMA)[None], len(q), axis=0))
    dr = min(1.0 / nx, 1.0 / 4)
    want = 0.2 * 0.5 * dr / (math.hypot(u, v) + math.sqrt(GAMMA * p / rho))
    assert o.time_step() == pytest.approx(want, rel=1e-14)
    for _ in range(9):
        o.time_step()
    assert o.time_step() == pytest.approx(0.5 * dr / (math.hypot(u, v) + math.sqrt(GAMMA * p / rho)), rel=1e-14)


def test_geometry_closed_forms():
    q, c = U.box(6, 5, size=2.0, code=U.REFLECTIVE)
    o = hfo.UnsOracle(q, c, [0, 0, 0, 0], GAMMA, 0.5)
    cen, vol = o.geometry()
    np.testing.assert_allclose(vol, (2.0 / 6) * (2.0 / 5), rtol=1e-14)
    np.testing.assert_allclose(cen, q.mean(axis=1), rtol=1e-14, atol=1e-15)
    q, c = U.box(7, 6, size=1.0, code=U.REFLECTIVE, skew=0.3)
    _, vol = hfo.UnsOracle(q, c, [0, 0, 0, 0], GAMMA, 0.5).geometry()
    assert vol.sum() == pytest.approx(1.0, rel=1e-14)


def test_connectivity_is_symmetric():
    q, c = U.box(7, 6, code=U.TRANSMISSIVE, skew=0.3)
    nb = hfo.UnsOracle(q, c, [0, 0, 0, 0], GAMMA, 0.5).connectivity()
    assert (nb[c != 0] == c[c != 0]).all()
    for cell in range(len(q)):
        for k in range(4):
            d = nb[cell, k]
            if d >= 0:
                assert cell in nb[d]
    # interior cell (i, j) of the row-major box: bottom, right, top, left
    i, j, nx = 3, 2, 7
    assert list(nb[j * nx + i]) == [(j - 1) * nx + i, j * nx + i + 1, (j + 1) * nx + i, j * nx + i - 1]


def test_hmsh_round_trip(tmp_path):
    q, c = U.ramp_channel(6, 4)
    p = str(tmp_path / "m.hmsh")
    U.write_hmsh(p, q, c)
    q2, c2 = U.read_hmsh(p)
    np.testing.assert_array_equal(q2, q)
    np.testing.assert_array_equal(c2, c)


def test_fma_is_correctly_rounded():
    from fractions import Fraction as F
    rng = np.random.default_rng(0)
    a = rng.standard_normal(3000) * 10.0 ** rng.integers(-4, 4, 3000)
    b = rng.standard_normal(3000) * 10.0 ** rng.integers(-4, 4, 3000)
    c = -a * b * (1 + rng.standard_normal(3000) * 1e-10)
    c[::3] = rng.standard_normal(1000)
    want = np.array([float(F(x) * F(y) + F(z)) for x, y, z in zip(a, b, c)])
    np.testing.assert_array_equal(U.fma(a, b, c), want)


@pytest.mark.skipif(not os.path.exists(KUMAR), reason="scramjet golden not generated")
def test_block_mesh_matches_reference_generator(tmp_path):
    g = json.load(open(KUMAR))
    q, c = U.block_mesh(g["mesh_blocks"])
    assert len(q) == g["ncells"]
    p = str(tmp_path / "kumar.hmsh")
    U.write_hmsh(p, q, c)
    text = open(p, "rb").read()
    assert len(text) == g["hmsh_bytes"]
    assert hashlib.sha256(text).hexdigest() == g["hmsh_sha256"]


def test_block_mesh_corners_and_codes():
    blocks = [{"coordinates": [[0.0, 0.0], [2.0, 0.5], [2.0, 1.5], [0.0, 1.0]], "cells": [4, 3],
               "boundary_conditions": [-1, -2, -1, -3]}]
    q, c = U.block_mesh(blocks)
    assert q.shape == (12, 4, 2)
    np.testing.assert_array_equal(q[0, 0], [0.0, 0.0])        # eps-major: cell 0 is (0, 0)
    np.testing.assert_array_equal(q[2, 3], [0.0, 1.0])        # cell 2 is (0, 2): top left corner
    np.testing.assert_array_equal(q[11, 2], [2.0, 1.5])
    np.testing.assert_array_equal(c[0], [-1, 0, 0, -3])
    np.testing.assert_array_equal(c[11], [0, -2, -1, 0])


@pytest.mark.skipif(not os.path.exists(KUMAR), reason="scramjet golden not generated")
def test_oracle_matches_reference_on_shipped_scramjet_case():
    g = json.load(open(KUMAR))
    q, c = U.block_mesh(g["mesh_blocks"])
    case = g["case"]
    gamma, cfl = case["model"]["gamma"], case["simulation"]["cfl"]
    o = hfo.UnsOracle(q, c, U.prim_to_cons(case["mesh"]["boundaries"]["inlet"], gamma), gamma, cfl)
    conn = np.ascontiguousarray(o.connectivity(), "<i4")
    assert hashlib.sha256(conn.tobytes()).hexdigest() == g["connectivity_sha256"]
    o.set_fields(U.initial_fields(q, case["initial_condition"], gamma))
    dts = o.run(g["nsteps"])
    assert [float(v).hex() for v in dts] == g["dts_hex"]
    out = np.ascontiguousarray(o.fields(), "<f8")
    assert hashlib.sha256(out.tobytes()).hexdigest() == g["fields_sha256"]
