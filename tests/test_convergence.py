"""Convergence to the exact solution (not a parity check): the shipped Toro
test 1 shock tube (cases/fvm_2d_cart/shock_tubes/toro/toro_shock_tube_01_*:
left (rho, u, p) = (1, 0.75, 1), right (0.125, 0, 0.1), interface x = 0.3,
t = 0.2, a sonic rarefaction, a contact and a shock) on 100, 200 and 400
cells, density L1 error against Toro's exact Riemann solution
(tests/exact_riemann.py, itself checked against Toro's tabulated star
states).  Pins the oracle's physics independently of the reference."""
import numpy as np
import pytest

import exact_riemann as ER
import hfo
from hyperflow_amd import case as C

LEFT, RIGHT, X0, T_END = (1.0, 0.75, 1.0), (0.125, 0.0, 0.1), 0.3, 0.2


def toro1(nx, scheme, limiter=None):
    cfg = {"model": {"model": "EulerIdealGas", "gamma": 1.4},
           "mesh": {"blocks": [{"extents": {"x_left": 0.0, "x_right": 1.0, "y_bottom": 0.0, "y_top": 0.05},
                                "boundaries": {"left": {"type": "Transmissive"}, "right": {"type": "Transmissive"},
                                               "bottom": {"type": "Reflective"}, "top": {"type": "Reflective"}},
                                "x_cells": nx, "y_cells": 4}]},
           "initial_condition": {"type": "QuadRiemannProblem", "x_interface": X0, "y_interface": 0.025,
                                 "variables": "Primitive", "init_ne_state": [0.125, 0.0, 0.0, 0.1],
                                 "init_nw_state": [1.0, 0.75, 0.0, 1.0], "init_sw_state": [1.0, 0.75, 0.0, 1.0],
                                 "init_se_state": [0.125, 0.0, 0.0, 0.1]},
           "scheme": {"type": scheme, "riemann": "HLLCEuler", "ode": "ForwardEuler"},
           "simulation": {"cfl": 0.5 if scheme == "MUSCLHancock" else 0.8, "t_start": 0.0, "t_end": T_END,
                          "animation": {"duration": 1.0, "fps": 30.0}}}
    if limiter:
        cfg["scheme"]["limiter"] = limiter
    return C.Case.from_json(cfg)


def density_error(nx, scheme, limiter=None):
    case = toro1(nx, scheme, limiter)
    m = hfo.OracleMesh(case)
    t = 0.0
    while t < T_END:
        t += m.step()
    g = case.ghost
    rho = m.fields(0)[0, g + 1, g:-g]              # one interior row (the solution is 1-D)
    x = (np.arange(nx) + 0.5) / nx
    exact = ER.sample(LEFT, RIGHT, x, t, X0)[:, 0]
    return float(np.mean(np.abs(rho - exact))), t


def test_exact_solver_matches_toro_tables():
    # Toro, 3rd ed., table 4.3: star pressure and velocity of tests 1-4
    for wl, wr, (p, u) in [((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), (0.30313, 0.92745)),
                           ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), (0.00189, 0.0)),
                           ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), (460.894, 19.5975)),
                           ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), (46.0950, -6.19633))]:
        ps, us = ER.star(wl, wr)
        assert ps == pytest.approx(p, rel=2e-5, abs=1e-5)
        assert us == pytest.approx(u, rel=2e-5, abs=1e-5)


@pytest.mark.parametrize("scheme,limiter,bound", [("Godunov", None, 0.008), ("MUSCLHancock", "VanLeer", 0.004),
                                                  ("MUSCLHancock", "MinBee", 0.0045)])
def test_toro1_converges_to_the_exact_solution(scheme, limiter, bound):
    errs = [density_error(n, scheme, limiter)[0] for n in (100, 200, 400)]
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[0] / errs[2] > 2.2, errs          # at least ~0.6 order in L1 across a shock and a contact
    assert errs[2] < bound, errs


def uns_density_error(nx):
    """The unstructured solver on a channel of nx x 2 quadrilaterals (reflective
    walls, transmissive ends), the same shock tube."""
    from hyperflow_amd import uns as U
    q, c = U.box(nx, 2, size=1.0, code=U.REFLECTIVE)
    q = q * np.array([1.0, 0.05])                 # [0, 1] x [0, 0.05]
    c[:, 1] = np.where(c[:, 1] != 0, U.TRANSMISSIVE, 0)
    c[:, 3] = np.where(c[:, 3] != 0, U.TRANSMISSIVE, 0)
    cen = U.centroids(q)
    left = U.prim_to_cons([1.0, 0.75, 0.0, 1.0], 1.4)
    right = U.prim_to_cons([0.125, 0.0, 0.0, 0.1], 1.4)
    f = np.where((cen[:, 0] <= X0)[:, None], left[None], right[None])
    o = hfo.UnsOracle(q, c, [0, 0, 0, 0], 1.4, 0.8)
    o.set_fields(f)
    t = 0.0
    while t < T_END:
        dt = o.time_step()
        o.advance(dt)
        t += dt
    rho = o.fields()[:nx, 0]                      # the bottom row of cells
    exact = ER.sample(LEFT, RIGHT, cen[:nx, 0], t, X0)[:, 0]
    return float(np.mean(np.abs(rho - exact)))


def test_unstructured_toro1_converges_to_the_exact_solution():
    errs = [uns_density_error(n) for n in (100, 200, 400)]
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[0] / errs[2] > 2.2, errs
    assert errs[2] < 0.008, errs
