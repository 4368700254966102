"""Writes the golden fixtures of tests/golden/ by running the reference's own
classes (oracle/_ref/ref_harness, built from the reference sources by
oracle/Makefile).  Run in a checkout where the reference is present:

    python tests/golden/make_golden.py

Each fixture is an .npz with the case JSON (as text), the initial padded
fields, the fields after `nsteps` steps and every TimeStep result.  Only data
is stored; no reference source.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import hfo  # noqa: E402
from hyperflow_amd import case as C  # noqa: E402


def blast_row(n, limiter):
    """Three blocks in a row; a pressure blast in block 0 whose wave speeds
    grow during the first steps, so the compat predictor dt of blocks 1 and 2
    differs from the one computed before block 0 is updated."""
    c = C.quad_case(3 * n, n, 0, scheme="MUSCLHancock", limiter=limiter, cfl=0.5,
                    extents=(0.0, 3.0, 0.0, 1.0))
    blocks = []
    for k in range(3):
        bnd = {"bottom": {"type": "Reflective"}, "top": {"type": "Reflective"},
               "left": {"type": "Transmissive"} if k == 0 else {"type": "Internal", "block": k - 1},
               "right": {"type": "Internal", "block": k + 1} if k < 2 else {"type": "Transmissive"}}
        blocks.append(C.Block(float(k), float(k + 1), 0.0, 1.0, n, n, bnd))
    c.blocks = blocks
    c.initial_condition = {"type": "SphericalRiemannProblem", "radius": 0.3, "x_origin": 0.5, "y_origin": 0.5,
                           "variables": "Primitive", "init_sphere_state": [1.0, 0.0, 0.0, 100.0],
                           "init_ext_state": [1.0, 0.0, 0.0, 1.0]}
    return c


def cases():
    out = []
    out.append(("godunov_quad_24x17_s0", C.quad_case(24, 17, 0, scheme="Godunov", cfl=0.8), None, (1, 100)))
    out.append(("muscl_minbee_quad_24x17_s1", C.quad_case(24, 17, 1, scheme="MUSCLHancock", limiter="MinBee",
                                                           cfl=0.5), None, (1, 100)))
    out.append(("muscl_vanleer_reflective_20x20_s2",
                C.quad_case(20, 20, 2, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5, bc="Reflective"),
                None, (1, 100)))
    c = C.quad_case(16, 12, 3, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
    out.append(("muscl_vanleer_perturbed_16x12_s3", c, C.perturbed_fields(c, 3), (1, 100)))
    out.append(("step_channel_minbee_8", C.step_channel(8, "MinBee", 0.5), None, (1, 100)))
    g = C.step_channel(6, "MinBee", 0.5)
    g.scheme, g.limiter = "Godunov", None
    out.append(("step_channel_godunov_6", g, None, (1, 100)))
    out.append(("blast_row_minbee_3x12", blast_row(12, "MinBee"), None, (1, 60)))
    out.append(("blast_row_vanleer_3x10", blast_row(10, "VanLeer"), None, (1, 60)))
    for p in ("shock_tubes/toro/toro_shock_tube_01_godunov_hllc_100x5.json",
              "shock_tubes/toro/toro_shock_tube_01_muscl_hancock_vanleer_hllc_100x5.json",
              "sphericalrp/sphericalrp_muscl_hancock_vanleer_hllc_50x50.json"):
        src = os.path.join("/root/reference/cases/fvm_2d_cart", p)
        if os.path.exists(src):
            out.append(("case_" + os.path.basename(p)[:-5], C.Case.load(src), None, (1, 40)))
    return out


def main():
    if not hfo.have_reference():
        sys.exit("oracle/_ref/ref_harness is not built (needs the reference checkout)")
    for name, case, fields, steps in cases():
        init = fields if fields is not None else case.initial_fields()
        data = {"case_json": np.array(json.dumps(case.to_json())), "nblocks": np.array(len(case.blocks))}
        for b, U in enumerate(init):
            data[f"init_{b}"] = U
        for n in steps:
            f, calls, sdt, _ = hfo.run_reference(case, n, init)
            for b, U in enumerate(f):
                data[f"after{n}_{b}"] = U
            data[f"calls{n}"] = calls
            data[f"stepdt{n}"] = sdt
        data["steps"] = np.array(steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
        print("wrote", name)


if __name__ == "__main__":
    main()


def cli_goldens():
    """Per-frame SHA-256 of the results_%06d.csv files the reference's own
    program (oracle/_ref/ref_cli) writes for shipped cases."""
    import hashlib
    import subprocess
    import tempfile
    ref_cli = os.path.join(ROOT, "oracle", "_ref", "ref_cli")
    out = {}
    for p in ("shock_tubes/toro/toro_shock_tube_01_godunov_hllc_100x5.json",
              "shock_tubes/toro/toro_shock_tube_01_muscl_hancock_vanleer_hllc_100x5.json"):
        src = os.path.join("/root/reference/cases/fvm_2d_cart", p)
        with tempfile.TemporaryDirectory() as td:
            case = C.Case.load(src)
            cj = os.path.join(td, "case.json")
            case.write_json(cj)
            r = subprocess.run([ref_cli, cj], cwd=td, stdout=subprocess.PIPE, text=True, check=True)
            steps = sum(1 for ln in r.stdout.splitlines() if ln.startswith("Steps:"))
            frames = sorted(f for f in os.listdir(td) if f.startswith("results_"))
            hashes = {f: hashlib.sha256(open(os.path.join(td, f), "rb").read()).hexdigest() for f in frames}
            first = open(os.path.join(td, frames[1])).read()
        out[os.path.basename(p)[:-5]] = {"case": case.to_json(), "progress_lines": steps, "frames": hashes,
                                         "frame_000001": first}
    with open(os.path.join(HERE, "cli_frames.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote cli_frames.json")


if __name__ == "__main__" and "--cli" in sys.argv:
    cli_goldens()
