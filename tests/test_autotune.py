"""The autotuner's problem classes and cache lookup (host logic only; the
tuning runs themselves are GPU tests)."""
import json

from hyperflow_amd import autotune
from hyperflow_amd import case as C


def c4(seed, count=512):
    return [C.quad_case(nx, ny, 1000 + k, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
            for k, (nx, ny) in enumerate(C.batched_sizes(count, seed))]


def test_classes_of_the_benchmark_workloads():
    k = autotune.class_key(autotune.problem_class(C.muscl_step_channel(1024), "fp64"))
    assert k == "MUSCLHancock/MinBee/fp64/compat/p1/b2-8/n4^10"
    k = autotune.class_key(autotune.problem_class(C.muscl_quadrants(0, 8192), "fp32"))
    assert k == "MUSCLHancock/VanLeer/fp32/compat/p1/b1/n4^13"
    k = autotune.class_key(autotune.problem_class(C.godunov_quadrants(0), "fp64"))
    assert k == "Godunov/None/fp64/compat/p1/b1/n4^9"
    # a new draw of the batched workload falls in the same class
    a = autotune.problem_class(c4(0), "fp32")
    b = autotune.problem_class(c4(7), "fp32")
    assert a == b and a["problems"] == "64+"


def _entry(cls, rows):
    return {"tile_rows": rows, "tail_rows": 6, "fix_rows": 0, "ms_per_step": 1.0, "spread": 0.01, "class": cls}


def test_lookup_exact_then_nearest_class(tmp_path, monkeypatch):
    cache = tmp_path / "tuning.json"
    monkeypatch.setattr(autotune, "CACHE", str(cache))
    base = autotune.problem_class(c4(0), "fp32")
    single = autotune.problem_class(C.muscl_quadrants(0, 8192), "fp32")
    other = dict(single, precision="fp64")
    cache.write_text(json.dumps({autotune.class_key(base): _entry(base, 48),
                                 autotune.class_key(single): _entry(single, 64),
                                 autotune.class_key(other): _entry(other, 16)}))
    # a C4 draw with a new seed: its own class is tuned
    assert autotune.cached_params(c4(3), "fp32") == {"tile_rows": 48, "tail_rows": 6, "fix_rows": 0}
    assert autotune.has_class(c4(3), "fp32")
    # an untuned single-block size: the nearest single-block fp32 class, never the fp64 entry
    assert not autotune.has_class(C.muscl_quadrants(0, 2048), "fp32")
    assert autotune.cached_rows(C.muscl_quadrants(0, 2048), "fp32") == 64
    assert autotune.cached_rows(C.muscl_quadrants(0, 2048), "fp64") == 16
    # nothing tuned for Godunov: library defaults
    assert autotune.cached_params(C.godunov_quadrants(0), "fp64") == {"tile_rows": 0, "tail_rows": 0, "fix_rows": 0}
