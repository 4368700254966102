"""Parity of the gfx950 unstructured solver (include/hyperflow_uns.h) with the
reference's src/fvm_2d_uns: bitwise equal connectivity, geometry, fields and
TimeStep results against the reference's goldens (tests/golden/uns_*.npz,
uns_kumar_coarse.json) and against the oracle on fresh inputs."""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import hfo
from gpu_util import say
from hyperflow_amd import uns as U

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "uns_*.npz")))
KUMAR = os.path.join(HERE, "golden", "uns_kumar_coarse.json")


def load(path):
    g = np.load(path)
    cj = json.loads(str(g["case_json"]))
    gamma, cfl = cj["model"]["gamma"], cj["simulation"]["cfl"]
    inlet = U.prim_to_cons(cj["mesh"]["boundaries"]["inlet"], gamma)
    return g, cj, inlet, gamma, cfl


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_device_run_matches_reference_golden(path):
    g, cj, inlet, gamma, cfl = load(path)
    n = int(g["steps"])
    say(f"uns golden {os.path.basename(path)}: {len(g['quads'])} cells, {n} steps")
    s = U.UnsSolver(g["quads"], g["codes"], inlet, gamma, cfl)
    assert s.bytes_per_cell() in (184, 248)   # 2 or 4 rotation terms per edge
    np.testing.assert_array_equal(s.connectivity(), g["connectivity"])
    cen, vol = s.geometry()
    np.testing.assert_array_equal(cen, g["centroids"])
    np.testing.assert_array_equal(vol, g["volumes"])
    s.upload(g["init"])
    np.testing.assert_array_equal(s.download(), g["init"])
    s.run(n)
    s.synchronize()
    np.testing.assert_array_equal(s.dt_history(), g["dts"])
    np.testing.assert_array_equal(s.download(), g["fields"])
    s.close()


@pytest.mark.parametrize("path", [p for p in GOLDEN if "nonphysical" not in p][:2] + [p for p in GOLDEN if "nonphysical" in p],
                         ids=lambda p: os.path.basename(p)[:-4])
def test_step_by_step_matches_fused_run(path):
    """time_step + advance per step (the Simulation::solve path, with room
    for host-side dt clamps) equals run()."""
    g, cj, inlet, gamma, cfl = load(path)
    say(f"uns step-by-step {os.path.basename(path)}")
    s = U.UnsSolver(g["quads"], g["codes"], inlet, gamma, cfl)
    s.upload(g["init"])
    dts = []
    for _ in range(int(g["steps"])):
        dt = s.time_step()
        s.advance(dt)
        dts.append(dt)
    np.testing.assert_array_equal(np.array(dts), g["dts"])
    np.testing.assert_array_equal(s.dt_history(), g["dts"])
    np.testing.assert_array_equal(s.download(), g["fields"])
    # mixing: run() continues from a step-by-step state and counter
    s.run(5)
    o = hfo.UnsOracle(g["quads"], g["codes"], inlet, gamma, cfl)
    o.set_fields(g["init"])
    odts = o.run(int(g["steps"]) + 5)
    np.testing.assert_array_equal(s.dt_history(), odts)
    np.testing.assert_array_equal(s.download(), o.fields())
    s.close()


@pytest.mark.parametrize("path", [p for p in GOLDEN if "nonphysical" not in p][:1], ids=lambda p: os.path.basename(p)[:-4])
def test_graph_replays_match_eager(path):
    """HIP-graph loops (cached per step count and starting state; the call
    counter on the device) equal eager steps bit for bit, over several
    replays of the same graphs with host work in between, odd counts that
    change the buffer, slot and call-count parities, the first ten calls'
    reduced CFL inside a graph, and a step-by-step interlude."""
    g, cj, inlet, gamma, cfl = load(path)
    say(f"uns graph vs eager {os.path.basename(path)}")
    plan = [3, 7, 3, 7, 6, 5, 3, 7]
    eager = U.UnsSolver(g["quads"], g["codes"], inlet, gamma, cfl)
    graph = U.UnsSolver(g["quads"], g["codes"], inlet, gamma, cfl)
    for s in (eager, graph):
        s.upload(g["init"])
    for k, n in enumerate(plan):
        eager.run(n)
        graph.run(n, graph=True)
        graph.synchronize()
        np.testing.assert_array_equal(graph.download(), eager.download())   # host work between replays
        if k == 4:   # a host-driven step in between
            for s in (eager, graph):
                s.advance(s.time_step())
    np.testing.assert_array_equal(graph.dt_history(), eager.dt_history())
    np.testing.assert_array_equal(graph.download(), eager.download())
    eager.close()
    graph.close()


def test_random_mesh_and_states_match_oracle():
    rng = np.random.default_rng(11)
    q, c = U.box(37, 29, code=U.TRANSMISSIVE, skew=0.35)
    c[::7, 1] = np.where(c[::7, 1] != 0, U.REFLECTIVE, 0)
    c[:37, 0] = U.INLET                   # the bottom row's floor
    W = np.stack([rng.uniform(0.3, 2.0, len(q)), rng.uniform(-1, 1, len(q)), rng.uniform(-1, 1, len(q)),
                  rng.uniform(0.2, 3.0, len(q))], axis=1)
    U0 = np.array([U.prim_to_cons(w, 1.4) for w in W])
    inlet = U.prim_to_cons([1.5, 0.8, 0.1, 1.2], 1.4)
    say(f"uns random {len(q)} cells")
    s = U.UnsSolver(q, c, inlet, 1.4, 0.6)
    o = hfo.UnsOracle(q, c, inlet, 1.4, 0.6)
    np.testing.assert_array_equal(s.connectivity(), o.connectivity())
    s.upload(U0)
    o.set_fields(U0)
    s.run(25)
    dts = o.run(25)
    np.testing.assert_array_equal(s.dt_history(), dts)
    np.testing.assert_array_equal(s.download(), o.fields())
    s.close()


def test_advance_with_given_dt_matches_oracle():
    g, cj, inlet, gamma, cfl = load(GOLDEN[0])
    s = U.UnsSolver(g["quads"], g["codes"], inlet, gamma, cfl)
    o = hfo.UnsOracle(g["quads"], g["codes"], inlet, gamma, cfl)
    s.upload(g["init"])
    o.set_fields(g["init"])
    for dt in (1e-4, 3.3e-4, 7e-5):
        s.advance(dt)
        o.advance(dt)
    np.testing.assert_array_equal(s.download(), o.fields())
    assert len(s.dt_history()) == 0
    s.close()


@pytest.mark.skipif(not os.path.exists(KUMAR), reason="scramjet golden not generated")
def test_shipped_scramjet_case_coarse_mesh():
    """scramjet_kumar_godunov_hllc.json on the mesh its coarse block file
    generates: connectivity, fields after the golden's steps and every dt
    equal the reference's."""
    k = json.load(open(KUMAR))
    q, c = U.block_mesh(k["mesh_blocks"])
    case = k["case"]
    gamma, cfl = case["model"]["gamma"], case["simulation"]["cfl"]
    inlet = U.prim_to_cons(case["mesh"]["boundaries"]["inlet"], gamma)
    say(f"uns scramjet coarse: {len(q)} cells, {k['nsteps']} steps")
    s = U.UnsSolver(q, c, inlet, gamma, cfl)
    conn = np.ascontiguousarray(s.connectivity(), "<i4")
    assert hashlib.sha256(conn.tobytes()).hexdigest() == k["connectivity_sha256"]
    s.upload(U.initial_fields(q, case["initial_condition"], gamma))
    s.run(k["nsteps"])
    s.synchronize()
    assert [float(v).hex() for v in s.dt_history()] == k["dts_hex"]
    out = np.ascontiguousarray(s.download(), "<f8")
    for i, row in zip(k["sample_cells"], k["sample_fields_hex"]):
        assert [float(v).hex() for v in out[i]] == row, f"cell {i}"
    assert hashlib.sha256(out.tobytes()).hexdigest() == k["fields_sha256"]
    s.close()


def test_bad_sizes_are_rejected():
    q, c = U.box(4, 4)
    s = U.UnsSolver(q, c)
    with pytest.raises(Exception, match="4\\*ncells"):
        U._check(U._lib().hf_uns_upload(s._h, np.zeros(5), 5), "hf_uns_upload")
    with pytest.raises(Exception, match="nsteps"):
        U._check(U._lib().hf_uns_run(s._h, -1, 0), "hf_uns_run")
    with pytest.raises(Exception, match="flags"):
        U._check(U._lib().hf_uns_run(s._h, 1, 2), "hf_uns_run")   # 1 = HF_UNS_RUN_GRAPH; 2 is unknown
    s.close()


CLI_GOLDEN = os.path.join(HERE, "golden", "uns_cli_frames.json")


@pytest.mark.parametrize("name", sorted(json.load(open(CLI_GOLDEN))) if os.path.exists(CLI_GOLDEN) else [])
def test_cli_frames_match_reference_program(name, tmp_path):
    """bin/hyperflow_uns writes the same results_%06d.csv frames, byte for
    byte, as the reference's unstructured program on the same case."""
    import subprocess
    g = json.load(open(CLI_GOLDEN))[name]
    U.write_hmsh(str(tmp_path / "mesh.hmsh"), np.array(g["quads"]), np.array(g["codes"], np.int32))
    with open(tmp_path / "case.json", "w") as f:
        json.dump(g["case"], f)
    exe = os.path.join(os.path.dirname(HERE), "hyperflow_amd", "bin", "hyperflow_uns")
    say(f"uns cli {name}")
    r = subprocess.run([exe, "case.json"], cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    assert sum(1 for ln in r.stdout.splitlines() if ln.startswith("Steps:")) == g["progress_lines"]
    frames = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("results_"))
    assert frames == sorted(g["frames"])
    for fr in frames:
        assert hashlib.sha256((tmp_path / fr).read_bytes()).hexdigest() == g["frames"][fr], fr
