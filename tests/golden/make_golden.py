"""Writes the golden fixtures of tests/golden/ by running the reference's own
classes (oracle/_ref/ref_harness, built from the reference sources by
oracle/Makefile).  Run in a checkout where the reference is present:

    python tests/golden/make_golden.py

Each fixture is an .npz with the case JSON (as text), the initial padded
fields, the fields after `nsteps` steps and every TimeStep result.  Only data
is stored; no reference source.

    python tests/golden/make_golden.py --nonphysical  # cells with negative pressure, zero density, NaN
    python tests/golden/make_golden.py --doublerp     # DoubleRiemannProblem ICs, Dirichlet in both variable kinds
    python tests/golden/make_golden.py --cli     # CSV frames of the reference program
    python tests/golden/make_golden.py --uns     # unstructured solver, small meshes
    python tests/golden/make_golden.py --uns-nonphysical  # the same with non-physical cells
    python tests/golden/make_golden.py --kumar   # shipped scramjet case, coarse mesh
    python tests/golden/make_golden.py --uns-cli # CSV frames of the reference's unstructured program
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


def doublerp_cases():
    """DoubleRiemannProblem initial conditions (doublerp.cpp:42-55, parsed at
    parser.cpp:325-356): three x-bands of constant state.  One case with
    "Primitive" variables whose left interface lies exactly on a cell
    centroid (the `<=` comparison decides that column), one with
    "Conserved" variables behind Dirichlet sides given in both variable kinds
    (parser.cpp:441-453)."""
    out = []
    nx, ny, g = 40, 16, 2
    dx = (1.0 - 0.0) / nx
    x_on_centroid = (0.0 - (float(g) - 0.5) * dx) + 13 * dx     # MeshBlock::calculate_centroids
    c = C.quad_case(nx, ny, 0, scheme="MUSCLHancock", limiter="MinBee", cfl=0.5, bc="Reflective")
    c.initial_condition = {"type": "DoubleRiemannProblem", "variables": "Primitive",
                           "x_lm_interface": x_on_centroid, "x_mr_interface": 0.7,
                           "init_left_state": [1.0, 0.75, -0.2, 1.0], "init_middle_state": [0.125, 0.0, 0.3, 0.1],
                           "init_right_state": [0.8, -0.5, 0.0, 0.6]}
    out.append(("doublerp_primitive_minbee_40x16", c, None, (1, 60)))
    c = C.quad_case(36, 20, 0, scheme="Godunov", cfl=0.8, bc="Reflective", extents=(-1.0, 2.0, 0.0, 1.5))
    c.initial_condition = {"type": "DoubleRiemannProblem", "variables": "Conserved",
                           "x_lm_interface": -0.25, "x_mr_interface": 1.1,
                           "init_left_state": [1.0, 0.5, 0.1, 2.6], "init_middle_state": [0.5, -0.1, 0.0, 1.2],
                           "init_right_state": [1.5, 0.0, 0.3, 3.9]}
    b = c.blocks[0].boundaries
    b["left"] = {"type": "Dirichlet", "variables": "Conserved", "flow_state": [1.2, 0.9, 0.0, 3.1]}
    b["right"] = {"type": "Dirichlet", "variables": "Primitive", "flow_state": [0.9, -0.3, 0.05, 1.3]}
    out.append(("doublerp_conserved_godunov_dirichlet_36x20", c, None, (1, 60)))
    return out


def nonphysical_cases():
    """Seeded cases with three non-physical cells (C.nonphysical_fields):
    negative pressure, zero density, a NaN momentum."""
    out = []
    for name, scheme, lim, seed in (("godunov", "Godunov", None, 4), ("minbee", "MUSCLHancock", "MinBee", 5),
                                    ("vanleer", "MUSCLHancock", "VanLeer", 6)):
        c = C.quad_case(20, 16, seed, scheme=scheme, limiter=lim, cfl=0.8 if lim is None else 0.5)
        out.append((f"nonphysical_{name}_20x16_s{seed}", c, C.nonphysical_fields(c, seed), (1, 5, 30)))
    c = C.step_channel(8, "MinBee", 0.5)
    out.append(("nonphysical_step_channel_minbee_8_s7", c, C.nonphysical_fields(c, 7, block=3), (1, 5, 30)))
    return out


def main(which=cases):
    if not hfo.have_reference():
        sys.exit("oracle/_ref/ref_harness is not built (needs the reference checkout)")
    for name, case, fields, steps in which():
        if fields is not None:
            init = fields
        else:
            # the reference's own initial fields (its IC classes, no steps);
            # the tests compare hyperflow_amd's initial fields against them
            init, _, _, _ = hfo.run_reference(case, 0)
            mine = case.initial_fields()
            if not all(np.array_equal(a, b) for a, b in zip(mine, init)):
                print(f"warning: {name}: initial fields differ from the reference's")
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


if __name__ == "__main__" and len(sys.argv) == 1:
    main()
if __name__ == "__main__" and sys.argv[1:] == ["--nonphysical"]:
    main(nonphysical_cases)
if __name__ == "__main__" and sys.argv[1:] == ["--doublerp"]:
    main(doublerp_cases)


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


# --------------------------------------------------- unstructured solver --
KUMAR = "/root/reference/cases/fvm_2d_uns/scramjet_kumar"


def uns_cases():
    """(name, quads, codes, case JSON without the mesh file, nsteps) of the
    unstructured goldens: a compression-ramp channel with inlet,
    transmissive and reflective walls; a sheared box with a blast
    (non-orthogonal cells, reflective walls); a transmissive box."""
    from hyperflow_amd import uns as U
    out = []
    q, c = U.ramp_channel(30, 12)
    inflow = [1.4, 3.0, 0.0, 1.0]
    out.append(("uns_ramp_channel_30x12", q, c,
                U.case_json("", (0.0, 3.0, 0.0, 1.0), inflow,
                            {"type": "Constant", "variables": "Primitive", "init_state": inflow}, cfl=0.5), 40))
    q, c = U.box(16, 16, code=U.REFLECTIVE, skew=0.3)
    out.append(("uns_skewed_box_blast_16x16", q, c,
                U.case_json("", (0.0, 1.0, 0.0, 1.0), [1.0, 0.0, 0.0, 1.0],
                            {"type": "SphericalRiemannProblem", "radius": 0.2, "x_origin": 0.4, "y_origin": 0.55,
                             "variables": "Primitive", "init_sphere_state": [1.0, 0.0, 0.0, 10.0],
                             "init_ext_state": [0.125, 0.0, 0.0, 0.1]}, cfl=0.8), 40))
    q, c = U.box(12, 10, code=U.TRANSMISSIVE)
    out.append(("uns_transmissive_box_12x10", q, c,
                U.case_json("", (0.0, 1.0, 0.0, 1.0), [1.0, 0.0, 0.0, 1.0],
                            {"type": "SphericalRiemannProblem", "radius": 0.3, "x_origin": 0.5, "y_origin": 0.5,
                             "variables": "Primitive", "init_sphere_state": [1.0, 0.5, -0.25, 2.0],
                             "init_ext_state": [0.5, 0.0, 0.0, 0.4]}, cfl=0.5), 30))
    # the loader's quirks: a boundary code on an edge that has a neighbour (the
    # neighbour wins) and an interior code on an outer edge (no neighbour: the
    # edge points at cell 0)
    q, c = U.box(5, 4, code=U.TRANSMISSIVE, skew=0.2)
    c[6, 1] = U.REFLECTIVE       # interior edge between cells 6 and 7
    c[12, 2] = U.INLET           # interior edge between cells 12 and 17
    c[9, 1] = 0                  # outer right edge of cell 9
    c[15, 3] = 0                 # outer left edge of cell 15
    out.append(("uns_loader_quirks_5x4", q, c,
                U.case_json("", (0.0, 1.0, 0.0, 1.0), [1.2, 0.4, 0.1, 1.1],
                            {"type": "SphericalRiemannProblem", "radius": 0.3, "x_origin": 0.5, "y_origin": 0.5,
                             "variables": "Primitive", "init_sphere_state": [1.0, 0.2, -0.1, 2.0],
                             "init_ext_state": [0.5, 0.0, 0.0, 0.5]}, cfl=0.5), 20))
    return out


def uns_nonphysical_cases():
    """The ramp channel started from its inflow state with three non-physical
    cells: negative total energy, zero density under momentum, NaN y-momentum."""
    from hyperflow_amd import uns as U
    q, c = U.ramp_channel(30, 12)
    inflow = [1.4, 3.0, 0.0, 1.0]

    def inject(f):
        f = f.copy()
        f[95, 3] = -0.5 * abs(f[95, 3])
        f[200, 0], f[200, 1] = 0.0, 0.3
        f[301, 2] = np.nan
        return f
    return [("uns_nonphysical_ramp_channel_30x12", q, c,
             U.case_json("", (0.0, 3.0, 0.0, 1.0), inflow,
                         {"type": "Constant", "variables": "Primitive", "init_state": inflow}, cfl=0.5), 30, inject)]


def uns_goldens(which=None):
    """Goldens of the reference's unstructured solver
    (oracle/_ref/ref_uns_harness): full arrays for small meshes."""
    import tempfile
    from hyperflow_amd import uns as U
    if not hfo.have_uns_reference():
        sys.exit("oracle/_ref/ref_uns_harness is not built (needs the reference checkout)")
    for name, q, c, cj, n, *inject in (which or uns_cases)():
        with tempfile.TemporaryDirectory() as td:
            mf = os.path.join(td, "mesh.hmsh")
            U.write_hmsh(mf, q, c)
            cj["mesh"]["file"] = mf
            r0 = hfo.run_uns_reference(cj, 0)
            start = inject[0](r0["fields"]) if inject else None
            if inject:
                r0["fields"] = start
            r = hfo.run_uns_reference(cj, n, fields=start)
        cj["mesh"]["file"] = "mesh.hmsh"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), case_json=np.array(json.dumps(cj)), quads=q,
                            codes=c, connectivity=r["connectivity"], centroids=r["centroids"],
                            volumes=r["volumes"], init=r0["fields"], fields=r["fields"], dts=r["dts"], steps=np.array(n))
        print("wrote", name)


def kumar_golden(nsteps=20):
    """The shipped scramjet case (scramjet_kumar_godunov_hllc.json) on the
    coarse mesh of scramjet_kumar_coarse_mesh.json: SHA-256 of the .hmsh
    text that the reference's mesh generator writes, of the reference
    solver's connectivity and of its fields after nsteps, plus the TimeStep
    results and a sample of cells.  The reference's mesh loader searches
    neighbours over all cell pairs, so this takes several minutes."""
    import hashlib
    import importlib.util
    import tempfile
    spec = importlib.util.spec_from_file_location(
        "generate_block_mesh", "/root/reference/src/mesh_generation/generate_block_mesh.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    cj = json.load(open(os.path.join(KUMAR, "scramjet_kumar_godunov_hllc.json")))
    with tempfile.TemporaryDirectory() as td:
        mf = os.path.join(td, "kumar.hmsh")
        gen.BlockStructuredMeshGenerator(os.path.join(KUMAR, "scramjet_kumar_coarse_mesh.json")).generate_mesh(mf)
        text = open(mf, "rb").read()
        cj["mesh"]["file"] = mf
        r = hfo.run_uns_reference(cj, nsteps)
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(len(r["fields"]), 64, replace=False))
    out = {"mesh_json": "scramjet_kumar_coarse_mesh.json", "case_json": "scramjet_kumar_godunov_hllc.json",
           "hmsh_sha256": hashlib.sha256(text).hexdigest(), "hmsh_bytes": len(text),
           "ncells": int(len(r["fields"])), "nsteps": nsteps,
           "connectivity_sha256": hashlib.sha256(np.ascontiguousarray(r["connectivity"], "<i4").tobytes()).hexdigest(),
           "fields_sha256": hashlib.sha256(np.ascontiguousarray(r["fields"], "<f8").tobytes()).hexdigest(),
           "dts_hex": [float(v).hex() for v in r["dts"]],
           "sample_cells": idx.tolist(), "sample_fields_hex": [[float(v).hex() for v in r["fields"][i]] for i in idx],
           "reference_seconds_per_step": r["elapsed"] / nsteps,
           "case": json.load(open(os.path.join(KUMAR, "scramjet_kumar_godunov_hllc.json"))),
           "mesh_blocks": json.load(open(os.path.join(KUMAR, "scramjet_kumar_coarse_mesh.json")))}
    with open(os.path.join(HERE, "uns_kumar_coarse.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote uns_kumar_coarse.json")


def uns_cli_goldens():
    """Per-frame SHA-256 of the results_%06d.csv files the reference's
    unstructured program (oracle/_ref/ref_uns_cli, its own main.cpp) writes
    for small cases, the mesh written next to the case file."""
    import hashlib
    import subprocess
    import tempfile
    from hyperflow_amd import uns as U
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_uns_cli")
    out = {}
    q, c = U.ramp_channel(30, 12)
    inflow = [1.4, 3.0, 0.0, 1.0]
    cases = [("ramp_channel_30x12", q, c,
              U.case_json("mesh.hmsh", (0.0, 3.0, 0.0, 1.0), inflow,
                          {"type": "Constant", "variables": "Primitive", "init_state": inflow}, cfl=0.5,
                          t_end=0.25)),
             ("skewed_box_blast_16x16",) + U.box(16, 16, code=U.REFLECTIVE, skew=0.3) + (
                 U.case_json("mesh.hmsh", (0.0, 1.0, 0.0, 1.0), [1.0, 0.0, 0.0, 1.0],
                             {"type": "SphericalRiemannProblem", "radius": 0.2, "x_origin": 0.4, "y_origin": 0.55,
                              "variables": "Primitive", "init_sphere_state": [1.0, 0.0, 0.0, 10.0],
                              "init_ext_state": [0.125, 0.0, 0.0, 0.1]}, cfl=0.8, t_end=0.1),)]
    for name, q, c, cj in cases:
        cj["simulation"]["animation"] = {"duration": 1.0, "fps": 10.0}
        with tempfile.TemporaryDirectory() as td:
            U.write_hmsh(os.path.join(td, "mesh.hmsh"), q, c)
            with open(os.path.join(td, "case.json"), "w") as f:
                json.dump(cj, f)
            r = subprocess.run([ref, "case.json"], cwd=td, stdout=subprocess.PIPE, text=True, check=True)
            steps = sum(1 for ln in r.stdout.splitlines() if ln.startswith("Steps:"))
            frames = sorted(f for f in os.listdir(td) if f.startswith("results_"))
            hashes = {f: hashlib.sha256(open(os.path.join(td, f), "rb").read()).hexdigest() for f in frames}
            last = open(os.path.join(td, frames[-1])).read()
        out[name] = {"case": cj, "quads": q.tolist(), "codes": c.tolist(), "progress_lines": steps,
                     "frames": hashes, "last_frame_head": "\n".join(last.splitlines()[:12])}
        print(name, steps, len(frames))
    with open(os.path.join(HERE, "uns_cli_frames.json"), "w") as f:
        json.dump(out, f, sort_keys=True)
    print("wrote uns_cli_frames.json")


if __name__ == "__main__" and "--uns-cli" in sys.argv:
    uns_cli_goldens()
if __name__ == "__main__" and "--uns-nonphysical" in sys.argv:
    uns_goldens(uns_nonphysical_cases)
if __name__ == "__main__" and "--uns" in sys.argv:
    uns_goldens()
if __name__ == "__main__" and "--kumar" in sys.argv:
    kumar_golden()
