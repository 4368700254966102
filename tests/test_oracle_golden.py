"""The CPU oracle against the golden vectors produced by the reference's own
classes (tests/golden/make_golden.py), and against the live reference
harness where it is built."""
import numpy as np
import pytest

import goldens
import hfo
from hyperflow_amd import case as C


@pytest.mark.parametrize("name", goldens.names())
def test_initial_fields_match_reference(name):
    g = goldens.load(name)
    if g["case"].initial_condition["type"] == "QuadRiemannProblem" and "perturbed" in name:
        mine = C.perturbed_fields(g["case"], 3)
    elif name.startswith("nonphysical_"):
        seed = int(name.rsplit("_s", 1)[1])
        mine = C.nonphysical_fields(g["case"], seed, block=3 if "step_channel" in name else 0)
    else:
        mine = g["case"].initial_fields()
    for a, b in zip(mine, g["init"]):
        assert goldens.same(a, b)


@pytest.mark.parametrize("name", goldens.names())
def test_oracle_matches_reference_bitwise(name):
    g = goldens.load(name)
    for n in g["steps"]:
        m = hfo.OracleMesh(g["case"], g["init"])
        stepdt, calls = m.run(n)
        assert goldens.same(stepdt, g[n]["stepdt"]), "step dt sequence"
        assert goldens.same(calls, g[n]["calls"]), "TimeStep call sequence"
        for b, ref in enumerate(g[n]["fields"]):
            assert goldens.same(m.fields(b), ref), f"block {b} after {n} steps"


@pytest.mark.skipif(not hfo.have_reference(), reason="reference harness not built")
@pytest.mark.parametrize("seed", [5, 6])
@pytest.mark.parametrize("kind", ["godunov", "minbee", "vanleer", "channel"])
def test_oracle_matches_live_reference(kind, seed):
    if kind == "godunov":
        case = C.quad_case(30, 22, seed, scheme="Godunov", cfl=0.8, bc="Reflective")
    elif kind == "channel":
        case = C.step_channel(10, "VanLeer" if seed % 2 else "MinBee", 0.5)
    else:
        case = C.quad_case(26, 31, seed, scheme="MUSCLHancock", limiter="MinBee" if kind == "minbee" else "VanLeer",
                           cfl=0.5)
    fields = case.initial_fields()
    rf, rc, rs, _ = hfo.run_reference(case, 60, fields)
    m = hfo.OracleMesh(case, fields)
    ds, dc = m.run(60)
    assert goldens.same(dc, rc)
    for b in range(len(rf)):
        assert goldens.same(m.fields(b), rf[b])
