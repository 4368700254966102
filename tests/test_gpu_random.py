"""Randomised multi-block parity: seeded block layouts (a bx x by grid of
blocks with random cell counts per block column and row, listed in a random
order), random outer boundary conditions (transmissive, reflective,
Dirichlet in primitive or conserved variables), both schemes and limiters,
perturbed quadrant data; the GPU path must equal the CPU oracle bit for bit
(fields incl. ghosts, every TimeStep result).  The block order changes the
ghost corners and the compat predictor dt (blocks update in list order)."""
import numpy as np
import pytest

import goldens
from gpu_util import gpu_run, oracle_run, say
from hyperflow_amd import case as C
from hyperflow_amd import native

pytestmark = pytest.mark.gpu


def random_case(seed: int, wide: bool = False) -> C.Case:
    """wide: block widths of 61..90 or 121..150 cells, so that the last
    60-column strip of a block carries 1..30 columns (a transposed band in the
    MUSCL launches), and heights up to 140 rows (several 60-row band items)."""
    rng = np.random.default_rng(seed)
    bx, by = int(rng.integers(1, 4)), int(rng.integers(1, 4))
    if wide:
        ncol = [int(rng.integers(1, 31)) + int(rng.choice([60, 120])) for _ in range(bx)]
        nrow = [int(rng.integers(6, 141)) for _ in range(by)]
    else:
        ncol = [int(rng.integers(6, 33)) for _ in range(bx)]
        nrow = [int(rng.integers(6, 33)) for _ in range(by)]
    order = rng.permutation(bx * by)              # list position of grid block (i, j)
    pos = {}
    for k, (j, i) in enumerate((j, i) for j in range(by) for i in range(bx)):
        pos[(i, j)] = int(order[k])

    def outer():
        kind = rng.choice(["Transmissive", "Reflective", "Dirichlet"])
        if kind != "Dirichlet":
            return {"type": str(kind)}
        w = [float(rng.uniform(0.3, 2.0)), float(rng.uniform(-0.5, 0.5)), float(rng.uniform(-0.5, 0.5)),
             float(rng.uniform(0.3, 2.0))]
        if rng.random() < 0.5:
            return {"type": "Dirichlet", "variables": "Primitive", "flow_state": w}
        return {"type": "Dirichlet", "variables": "Conserved", "flow_state": C.prim_to_cons(1.4, w)}

    sides = {"bottom": outer(), "right": outer(), "top": outer(), "left": outer()}
    blocks = [None] * (bx * by)
    for j in range(by):
        for i in range(bx):
            bnd = {"bottom": sides["bottom"] if j == 0 else {"type": "Internal", "block": pos[(i, j - 1)]},
                   "top": sides["top"] if j == by - 1 else {"type": "Internal", "block": pos[(i, j + 1)]},
                   "left": sides["left"] if i == 0 else {"type": "Internal", "block": pos[(i - 1, j)]},
                   "right": sides["right"] if i == bx - 1 else {"type": "Internal", "block": pos[(i + 1, j)]}}
            blocks[pos[(i, j)]] = C.Block(i / bx, (i + 1) / bx, j / by, (j + 1) / by, ncol[i], nrow[j], bnd)
    scheme = "Godunov" if rng.random() < 0.3 and not wide else "MUSCLHancock"
    limiter = None if scheme == "Godunov" else str(rng.choice(["MinBee", "VanLeer"]))
    base = C.quad_case(8, 8, seed, scheme=scheme, limiter=limiter, cfl=0.8 if scheme == "Godunov" else 0.5)
    base.blocks = blocks
    return base


@pytest.mark.parametrize("seed", range(16))
def test_random_multiblock_matches_oracle(seed):
    case = random_case(seed)
    f = C.perturbed_fields(case, seed)
    say(f"random case {seed}: {len(case.blocks)} blocks, {case.scheme} {case.limiter}")
    _, out, hist = gpu_run(case, f, 40)
    ref, calls, _ = oracle_run(case, f, 40)
    assert goldens.same(hist, calls), "TimeStep sequence"
    for b in range(len(case.blocks)):
        assert goldens.same(out[b], ref[b]), f"block {b}"


@pytest.mark.parametrize("seed", range(100, 108))
def test_random_wide_blocks_match_oracle(seed):
    """Blocks whose right edge is computed by transposed band items."""
    case = random_case(seed, wide=True)
    f = C.perturbed_fields(case, seed)
    say(f"random wide case {seed}: {[(b.nx, b.ny) for b in case.blocks]}, {case.limiter}")
    _, out, hist = gpu_run(case, f, 20)
    ref, calls, _ = oracle_run(case, f, 20)
    assert goldens.same(hist, calls), "TimeStep sequence"
    for b in range(len(case.blocks)):
        assert goldens.same(out[b], ref[b]), f"block {b}"


def _muscl_multiblock(seed):
    """A random layout of at least two blocks, MUSCL-Hancock (the compat
    predictor dt, so the guess and check launches run)."""
    for k in range(50):
        case = random_case(seed + 1000 * k)
        if case.scheme == "MUSCLHancock" and len(case.blocks) >= 2:
            return case, seed + 1000 * k
    raise RuntimeError("no multi-block MUSCL layout")


def test_batched_multiblock_problems_match_oracle():
    """Several multi-block problems in one context: the fix-up plans walk
    every problem.  Each problem must equal its own oracle run bit for bit,
    and the fix-up launches must have had work."""
    picks = [_muscl_multiblock(s) for s in (201, 202, 203, 204)]
    cases = [c for c, _ in picks]
    for c in cases:
        c.limiter = "VanLeer"   # one context: one scheme and limiter
    fields = [C.perturbed_fields(c, s) for c, s in picks]
    say(f"batched multi-block problems: {[len(c.blocks) for c in cases]} blocks")
    s = native.Solver(cases)
    try:
        s.upload_all([U for f in fields for U in f])
        s.run(30)
        s.synchronize()
        out = s.download_all()
        stats = s.fixup_stats()
    finally:
        s.close()
    say(f"  fix-up tallies {stats}")
    assert stats["guess"] + stats["check"] > 0
    first = 0
    for k, c in enumerate(cases):
        ref, _, _ = oracle_run(c, fields[k], 30)
        for b in range(len(c.blocks)):
            assert goldens.same(out[first + b], ref[b]), f"problem {k} block {b}"
        first += len(c.blocks)


def test_problem_of_more_than_64_blocks_matches_oracle():
    """A 10 x 7 grid of small blocks in one problem: the fix-up plan scans the
    blocks 64 at a time (two chunks) and the verdicts stride over them."""
    rng = np.random.default_rng(7)
    bx, by = 10, 7
    ncol = [int(rng.integers(6, 11)) for _ in range(bx)]
    nrow = [int(rng.integers(6, 11)) for _ in range(by)]
    order = rng.permutation(bx * by)
    pos = {(i, j): int(order[j * bx + i]) for j in range(by) for i in range(bx)}
    blocks = [None] * (bx * by)
    for j in range(by):
        for i in range(bx):
            bnd = {"bottom": {"type": "Reflective"} if j == 0 else {"type": "Internal", "block": pos[(i, j - 1)]},
                   "top": {"type": "Transmissive"} if j == by - 1 else {"type": "Internal", "block": pos[(i, j + 1)]},
                   "left": {"type": "Transmissive"} if i == 0 else {"type": "Internal", "block": pos[(i - 1, j)]},
                   "right": {"type": "Transmissive"} if i == bx - 1 else {"type": "Internal", "block": pos[(i + 1, j)]}}
            blocks[pos[(i, j)]] = C.Block(i / bx, (i + 1) / bx, j / by, (j + 1) / by, ncol[i], nrow[j], bnd)
    case = C.quad_case(8, 8, 7, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
    case.blocks = blocks
    f = C.perturbed_fields(case, 7)
    say(f"70-block problem, {sum(b.nx * b.ny for b in blocks)} cells")
    s, out, hist = gpu_run(case, f, 25)
    stats = s.fixup_stats()
    s.close()
    say(f"  fix-up tallies {stats}")
    ref, calls, _ = oracle_run(case, f, 25)
    assert goldens.same(hist, calls), "TimeStep sequence"
    for b in range(len(blocks)):
        assert goldens.same(out[b], ref[b]), f"block {b}"
