"""The `hyperflow` command-line driver: Simulation::solve semantics (t_end and
frame clamps, the duplicated first frame) and the reference's CSV format.
GPU test: every results_%06d.csv frame is byte-identical to what the
reference's own program writes (SHA-256 of each frame in
tests/golden/cli_frames.json).  CPU tests: errors raised before any device
work."""
import hashlib
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "hyperflow_amd", "bin", "hyperflow")
GOLD = os.path.join(ROOT, "tests", "golden", "cli_frames.json")


def golden():
    with open(GOLD) as f:
        return json.load(f)


def run_cli(case_json, workdir, *extra):
    path = os.path.join(workdir, "case.json")
    with open(path, "w") as f:
        json.dump(case_json, f, indent=1)
    return subprocess.run([CLI, path, "--output-dir", str(workdir)] + list(extra), stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=300)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(golden()))
def test_cli_frames_are_byte_identical_to_reference(name, tmp_path):
    g = golden()[name]
    print(f"cli {name}", flush=True)
    r = run_cli(g["case"], tmp_path)
    assert r.returncode == 0, r.stderr
    progress = sum(1 for ln in r.stdout.splitlines() if ln.startswith("Steps:"))
    assert progress == g["progress_lines"]
    frames = sorted(f for f in os.listdir(tmp_path) if f.startswith("results_"))
    assert frames == sorted(g["frames"])
    with open(os.path.join(tmp_path, "results_000001.csv")) as f:
        assert f.read() == g["frame_000001"]
    for fr in frames:
        with open(os.path.join(tmp_path, fr), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == g["frames"][fr], fr


def _case():
    return json.loads(json.dumps(golden()["toro_shock_tube_01_godunov_hllc_100x5"]["case"]))


def test_cli_rejects_malformed_json(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"model": {"model": "EulerIdealGas", "gamma": 1.4,,}}')
    r = subprocess.run([CLI, str(p)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 1 and "JSON parse error" in r.stderr


def test_cli_reports_missing_keys(tmp_path):
    c = _case()
    del c["simulation"]["cfl"]
    r = run_cli(c, tmp_path)
    assert r.returncode == 1 and "missing key 'cfl'" in r.stderr


@pytest.mark.parametrize("section,key,value,message", [
    ("scheme", "type", "WENO", "Scheme type 'WENO' not supported. Returning null scheme."),
    ("scheme", "ode", "RK4", "ODE Solver type 'RK4' not supported. Returning null ODE solver."),
    ("model", "model", "Navier", "Model type 'Navier' not supported. Returning null model."),
])
def test_cli_unknown_types_print_reference_messages(tmp_path, section, key, value, message):
    c = _case()
    c[section][key] = value
    r = run_cli(c, tmp_path)
    assert message in r.stdout
    assert r.returncode == 1


def test_cli_usage(tmp_path):
    r = subprocess.run([CLI], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
