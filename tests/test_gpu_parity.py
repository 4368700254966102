"""Parity of the gfx950 FP64 path with the reference (golden vectors) and the
CPU oracle.  Integer-exact expectations: the FP64 kernels evaluate the
reference's formulas in its order, so results must be identical (sign of zero
aside); the FP64 accuracy bar (rel L-inf <= 1e-12 after 1 step, <= 1e-9
after 100) is asserted on top."""
import numpy as np
import pytest

import goldens
import hfo
from gpu_util import gpu_run, interior, oracle_run, rel_linf, say
from hyperflow_amd import case as C
from hyperflow_amd import native

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", goldens.names())
def test_fused_run_matches_reference_golden(name):
    g = goldens.load(name)
    for n in g["steps"]:
        say(f"golden {name}: {n} steps")
        s, out, hist = gpu_run(g["case"], g["init"], n)
        assert goldens.same(hist, g[n]["calls"]), "TimeStep call sequence"
        for b, ref in enumerate(g[n]["fields"]):
            assert goldens.same(out[b], ref), f"block {b} after {n} steps"
        s.close()


@pytest.mark.parametrize("name", ["step_channel_minbee_8", "muscl_vanleer_perturbed_16x12_s3",
                                  "godunov_quad_24x17_s0"])
def test_graph_capture_matches_eager(name):
    g = goldens.load(name)
    say(f"graph {name}")
    s = native.Solver(g["case"])
    s.upload_all(g["init"])
    s.run(50, graph=True)
    s.run(50, graph=True)     # replay of the cached graph
    s.synchronize()
    out = s.download_all()
    for b, ref in enumerate(g[100]["fields"]):
        assert goldens.same(out[b], ref)
    assert goldens.same(s.dt_history(), g[100]["calls"])


@pytest.mark.parametrize("rows", [1, 3, 16, 64])
def test_tile_rows_do_not_change_results(rows):
    case = C.quad_case(130, 77, 9, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
    f = case.initial_fields()
    say(f"tile rows {rows}")
    _, out, _ = gpu_run(case, f, 20, tile_rows=rows)
    ref, _, _ = oracle_run(case, f, 20)
    assert goldens.same(out[0], ref[0])


@pytest.mark.parametrize("rows", [1, 2, 5])
@pytest.mark.parametrize("kind", ["godunov", "channel"])
def test_small_tile_rows_match_oracle(kind, rows):
    """Items of a few rows: Godunov on one block, and the 8-block channel whose
    compat dt needs fix-up launches (their items are capped at the tile rows)."""
    if kind == "godunov":
        case = C.quad_case(97, 61, 12, scheme="Godunov", cfl=0.8)
    else:
        case = C.step_channel(18, "MinBee", 0.5)
    f = case.initial_fields()
    say(f"small tile rows: {kind} rows {rows}")
    _, out, hist = gpu_run(case, f, 15, tile_rows=rows)
    ref, calls, _ = oracle_run(case, f, 15)
    for b in range(len(ref)):
        assert goldens.same(out[b], ref[b]), f"block {b}"
    assert goldens.same(hist, calls)


@pytest.mark.parametrize("scheme,lim", [("Godunov", None), ("MUSCLHancock", "MinBee"), ("MUSCLHancock", "VanLeer")])
def test_class_level_api_matches_oracle(scheme, lim):
    """Mesh::apply_bcs, TimeStep, Scheme::operator() and the per-block
    ForwardEuler loop, called one by one like Simulation::solve does."""
    case = C.step_channel(20, lim or "MinBee", 0.5)
    case.scheme, case.limiter = scheme, lim
    f = case.initial_fields()
    # disturb the constant state so that every block evolves differently
    rng = np.random.default_rng(3)
    for U in f:
        g = case.ghost
        U[0, g:-g, g:-g] *= 1.0 + 0.05 * rng.random(U[0, g:-g, g:-g].shape)
    say(f"class-level {scheme} {lim}")
    s = native.Solver(case)
    s.upload_all(f)
    m = hfo.OracleMesh(case, f)
    for step in range(12):
        s.apply_bcs()
        m.apply_bcs()
        for b in range(len(case.blocks)):
            assert goldens.same(s.download(b), m.fields(b)), f"ghosts of block {b}, step {step}"
        dt = s.time_step(0)
        assert dt == m.time_step()
        if step == 5:
            for b in range(len(case.blocks)):
                assert goldens.same(s.residual(b), m.residual(b)), f"residual of block {b}"
        s.advance([dt])
        for b in range(len(case.blocks)):
            Lb = m.residual(b)
            m.set_fields(b, dt * Lb + m.fields(b))
        for b in range(len(case.blocks)):
            assert goldens.same(s.download(b), m.fields(b)), f"block {b} after step {step}"
    assert s.calls(0) == m.calls


def test_config1_full_size_against_oracle():
    """Godunov + HLLC 400x400, transmissive, CFL 0.8, 100 steps."""
    for seed in (0, 1):
        case = C.godunov_quadrants(seed)
        f = case.initial_fields()
        for n in (1, 100):
            say(f"config1 seed {seed}: {n} steps")
            _, out, hist = gpu_run(case, f, n)
            ref, calls, _ = oracle_run(case, f, n)
            err = rel_linf(out, ref)
            assert err <= (1e-12 if n == 1 else 1e-9), err
            assert goldens.same(out[0], ref[0])
            assert goldens.same(hist, calls)


def test_step_channel_reduced_against_oracle():
    """The config-2 geometry (8 blocks, Dirichlet/Reflective/Internal/
    Transmissive, MUSCL MinBee, compat dt) at 64x64 per block, 500 steps."""
    case = C.muscl_step_channel(64)
    f = case.initial_fields()
    say("step channel 64: 500 steps")
    _, out, hist = gpu_run(case, f, 500)
    ref, calls, _ = oracle_run(case, f, 500)
    assert goldens.same(hist, calls)
    for b in range(8):
        assert goldens.same(out[b], ref[b]), f"block {b}"


def test_config2_full_size_first_steps_against_oracle():
    case = C.muscl_step_channel(1024)
    f = case.initial_fields()
    say("config2 full size: 3 steps")
    _, out, hist = gpu_run(case, f, 3)
    ref, calls, _ = oracle_run(case, f, 3, threads=16)
    assert goldens.same(hist, calls)
    for b in range(8):
        assert goldens.same(out[b], ref[b]), f"block {b}"


def test_uniform_mode_equals_compat_for_single_block():
    case = C.quad_case(90, 70, 4, scheme="MUSCLHancock", limiter="MinBee", cfl=0.5)
    f = case.initial_fields()
    say("uniform vs compat")
    _, a, ha = gpu_run(case, f, 30, dt_mode="compat")
    _, b, hb = gpu_run(case, f, 30, dt_mode="uniform")
    assert goldens.same(a[0], b[0]) and goldens.same(ha, hb)


def test_batched_problems_match_individual_runs():
    cases = [C.quad_case(40 + 7 * k, 33 + 5 * k, 20 + k, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
             for k in range(5)]
    fields = [c.initial_fields()[0] for c in cases]
    say("batched 5 problems")
    s, out, _ = gpu_run(cases, fields, 25)
    for k, c in enumerate(cases):
        ref, _, _ = oracle_run(c, [fields[k]], 25)
        assert goldens.same(out[k], ref[0]), f"problem {k}"


def test_conservation_sums_godunov_reflective_box():
    case = C.quad_case(64, 48, 2, scheme="Godunov", cfl=0.4, bc="Reflective")
    say("conservation")
    s = native.Solver(case)
    s.upload_all(case.initial_fields())
    s0 = s.conservation(0)
    s.run(200)
    s1 = s.conservation(0)
    np.testing.assert_allclose(s1[[0, 3]], s0[[0, 3]], rtol=1e-13)
    exact = np.array([np.sum(np.sort(U[k, 1:-1, 1:-1].ravel())) for U in [s.download(0)] for k in range(4)])
    np.testing.assert_allclose(s1, exact, rtol=1e-14, atol=1e-12)


def test_sampled_launch_timing_counts():
    """hf_profile(k): the update launches of steps 0, k, 2k, ... are timed."""
    g = goldens.load("step_channel_minbee_8")
    s = native.Solver(g["case"])
    s.upload_all(g["init"])
    say("profile sampling")
    for every, steps in ((1, 7), (3, 10), (4, 8)):
        s.profile(True, every=every)
        s.kernel_stats()
        s.run(steps)
        s.synchronize()
        launches, ms = s.kernel_stats()
        # one update launch per step: the main list and the fix-up rounds together
        assert launches == -(-steps // every), (every, steps, launches)
        assert ms > 0.0
    s.profile(False)
    s.run(3)
    assert s.kernel_stats() == (0, 0.0)
    s.close()
