"""bench.py's command line and the committed measurements it reads (no GPU):
every workload name resolves, the update-kernel workloads find their HBM
traffic figures at the default sizes, and more than one GPU is refused."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

# workload name and precision as bench.py's default run of each uses them
TRAFFIC_KEYS = [("step-channel", "fp64"), ("muscl-quadrants", "fp64"), ("muscl-quadrants", "fp32"),
                ("godunov-quadrants", "fp64")]


@pytest.mark.parametrize("workload,precision", TRAFFIC_KEYS)
def test_traffic_figures_exist_for_the_update_workloads(workload, precision):
    t = bench.traffic_from_profiles(f"{workload}_{precision}")
    assert isinstance(t, int) and t > 0


def test_traffic_figure_exists_for_the_unstructured_workload():
    assert bench.traffic_from_profiles("unstructured_r8") > 0


def test_traffic_keys_are_workload_names():
    with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
        keys = [k for k in json.load(f) if not k.startswith("_")]
    names = {"step-channel", "muscl-quadrants", "godunov-quadrants", "batch", "learned", "unstructured"}
    for k in keys:
        assert k.rsplit("_", 1)[0] in names, k


@pytest.mark.parametrize("name", ["godunov-quadrants", "step-channel", "muscl-quadrants", "batch"])
def test_workloads_build(name):
    cases, fields, text = bench.build_workload(name, 16, "fp64")
    assert len(cases) == len(fields) or name == "step-channel"
    assert text


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)


def test_valu_bound_of_the_fp64_muscl_loop():
    """About 23.4 Gcell-updates/s: 1 337 full-rate and 60 quarter-rate
    instructions per 60-cell wave-row on 1 024 SIMDs at 2.4 GHz."""
    v = bench.valu_bound("MUSCLHancock", "MinBee", "fp64", 18.0e9)
    assert 23300 < v["bound"] < 23450
    assert abs(v["frac"] - 18000.0 / v["bound"]) < 1e-4
    assert bench.valu_bound("Godunov", None, "fp64", 1e9) is None


def test_learned_check_seeds_took_no_part_in_training_or_choice():
    """The headline line's learned sub-object runs on seeds that neither the
    training (seeds 0..39, the held-out ones included) nor the choice of the
    network and its settings used."""
    import numpy as np
    z = np.load(os.path.join(ROOT, "hyperflow_amd", "learned_weights.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    lo, hi = meta["train_seeds"]
    used = set(range(lo, hi + 1)) | set(meta["holdout_seeds"])
    for stage in meta.get("unrolled", []):
        a, b = stage["train_seeds"]
        used |= set(range(a, b + 1)) | set(stage["holdout_seeds"])
    check, choice = set(bench.LEARNED_SEEDS_CHECK), set(bench.LEARNED_SEEDS_CHOICE)
    assert check and not check & choice
    assert not (check | choice) & used
    assert meta["every"] == meta["trained_every"]      # runs correct at the interval trained for
