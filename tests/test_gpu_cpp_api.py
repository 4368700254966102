"""The reference-typed C++ surface of include/hyperflow.hpp, driven by a copy
of the reference's per-block step loop (simulation.cpp:106-110:
std::bind(&Scheme::operator(), scheme, block) into
ODESolver::operator()(func, MeshBlock::get_fields(), dt), then
MeshBlock::set_fields) compiled against the header (tests/cpp/ref_loop.cpp).
Every field after the golden step counts equals the golden written by the
reference's own classes, bitwise, and so does every step dt; the goldens
without given fields start from the C++ initial-condition classes."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import goldens
import hfo
from gpu_util import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "bin", "ref_loop")
# goldens whose initial fields come from the case's initial condition (the
# others start from given fields: perturbed or non-physical cells)
FROM_IC = [n for n in goldens.names() if not (n.startswith("nonphysical_") or "perturbed" in n)]


def run_driver(case, nsteps, fields=None):
    with tempfile.TemporaryDirectory() as td:
        cj = os.path.join(td, "case.json")
        out = os.path.join(td, "out.bin")
        case.write_json(cj)
        args = [DRIVER, cj, str(nsteps), out]
        if fields is not None:
            fin = os.path.join(td, "fields.bin")
            with open(fin, "wb") as f:
                for U in fields:
                    f.write(np.ascontiguousarray(U, np.float64).tobytes())
            args.append(fin)
        r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        return hfo.read_ref_output(out)


@pytest.mark.parametrize("name", goldens.names())
def test_reference_loop_through_cpp_header_matches_golden(name):
    g = goldens.load(name)
    given = None if name in FROM_IC else g["init"]
    for n in g["steps"]:
        say(f"c++ reference loop {name}: {n} steps")
        fields, _, stepdt, _ = run_driver(g["case"], n, given)
        assert goldens.same(stepdt, g[n]["stepdt"]), "step dt sequence"
        for b, ref in enumerate(g[n]["fields"]):
            assert goldens.same(fields[b], ref), f"block {b} after {n} steps"
