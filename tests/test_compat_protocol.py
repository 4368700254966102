"""The reference-compatible predictor dt of the MUSCL step, met by three
launches (main list, guess, check; hf_kernels.h k_muscl / k_muscl_guess /
k_muscl_check), against the reference's sequential definition
(simulation.cpp:106-110 with muscl_hancock.cpp:246: block l's predictor dt
sees blocks 0..l-1 at U^{n+1}).

A model of the protocol at the level of per-block dt terms: old_c is block
c's term at U^n, and computing block c with a dt gives a new term f(c, dt)
(a random but fixed function: the protocol must not depend on what it is).
The model follows the kernels' decisions -- skip flags, the guess plan's
first block and minimum, the check's first wrong block and its rounds -- and
every block must end with the reference's dt and new term.
"""
import random

INF = float("inf")


def cfl_eff(cfl, k):
    return 0.2 * cfl if k <= 10 else cfl   # timestep.cpp:60-65


def reference(nb, o, f, kstep, cfl):
    dts, new = [], []
    for l in range(nb):
        dts.append(cfl_eff(cfl, kstep + 1 + l) * min([1e16] + new[:l] + o[l:]))
        new.append(f(l, dts[-1]))
    return dts, new


def three_launches(nb, o, f, skip, kstep, cfl, plant=False):
    """Returns (dt used per block, final new term per block, blocks computed
    by the guess launch, blocks recomputed by the check launch)."""
    cf = [cfl_eff(cfl, kstep + 1 + l) for l in range(nb)]
    O = min([1e16] + o)
    spec = [cf[l] * O for l in range(nb)]

    def exact(l, pre):
        return cf[l] * min([1e16] + [pre(c) for c in range(l)] + o[l:])

    # main launch: every block not skipped, speculative dt
    main = [None if skip[l] else f(l, spec[l]) for l in range(nb)]
    # guess launch: first wrong block L from the main list's terms, then every
    # block l >= L that needs it with cf_l * min(1e16, m_L)
    L = next((l for l in range(1, nb) if skip[l] or spec[l] != exact(l, lambda c: main[c])), nb)
    m = min([main[c] for c in range(L)] + o[L:]) if L < nb else 0.0
    if plant:   # a planted mistake: the guess ignores the updated blocks before L
        m = min(o[L:]) if L < nb else 0.0
    guess = {}
    for l in range(L, nb):
        g = cf[l] * min(1e16, m)
        if skip[l] or spec[l] != g:
            guess[l] = (g, f(l, g))
    if plant:   # ... and the check launch trusts the guess
        used = [guess[l][0] if l in guess else spec[l] for l in range(nb)]
        return used, [guess[l][1] if l in guess else main[l] for l in range(nb)], sorted(guess), []
    # check launch: the first block whose dt so far is not the exact one, then
    # rounds in block order
    used = [guess[l][0] if l in guess else (None if skip[l] else spec[l]) for l in range(nb)]
    final = [guess[l][1] if l in guess else main[l] for l in range(nb)]
    rechecked = []
    W = next((l for l in range(1, nb) if used[l] != exact(l, lambda c: final[c])), nb)
    for l in range(W, nb):
        ex = exact(l, lambda c: final[c])
        if used[l] != ex:
            used[l] = ex
            final[l] = f(l, ex)
            rechecked.append(l)
    used[0] = spec[0]
    return used, final, sorted(guess), rechecked


def random_case(seed):
    R = random.Random(seed)
    nb = R.randint(2, 150)
    vals = [R.choice([0.5, 1.0, 1.5, 2.0, 2.5, 3.0]) for _ in range(8)]
    o = [R.choice(vals) if R.random() < 0.9 else INF for _ in range(nb)]
    table = {}

    def f(l, dt):
        key = (l, dt)
        if key not in table:
            table[key] = R.choice(vals + [INF])
        return table[key]

    skip = [0] + [1 if R.random() < 0.3 else 0 for _ in range(nb - 1)]
    return nb, o, f, skip, R.choice([1, 5, 9, 10, 11, 50])


def test_three_launches_give_the_sequential_dts():
    checked = 0
    for seed in range(3000):
        nb, o, f, skip, kstep = random_case(seed)
        ref_dt, ref_new = reference(nb, o, f, kstep, 0.5)
        used, final, _, _ = three_launches(nb, o, f, skip, kstep, 0.5)
        assert used == ref_dt and final == ref_new, seed
        checked += 1
    assert checked == 3000


def test_check_launch_rarely_needed_and_guess_exact_in_the_common_pattern():
    """Block 3 sets the minimum and its new state is faster: every later block
    takes block 3's new term, which the guess launch finds in one pass."""
    nb = 8
    o = [2.0, 2.0, 2.0, 1.0, 2.0, 2.0, 2.0, 2.0]
    f = lambda l, dt: 0.9 if l == 3 else 1.8
    used, final, guessed, rechecked = three_launches(nb, o, f, [0] * nb, 50, 0.5)
    ref_dt, ref_new = reference(nb, o, f, 50, 0.5)
    assert used == ref_dt and final == ref_new
    assert guessed == [4, 5, 6, 7] and rechecked == []


def test_a_plan_that_misses_a_new_minimum_is_repaired_by_the_check_launch():
    """Block 5's recomputed state sets a new minimum the guess could not know:
    blocks 6 and 7 are recomputed in the check launch's rounds."""
    nb = 8
    o = [2.0, 2.0, 2.0, 1.0, 2.0, 2.0, 2.0, 2.0]
    f = lambda l, dt: {3: 0.9, 5: 0.5}.get(l, 1.8)
    used, final, guessed, rechecked = three_launches(nb, o, f, [0] * nb, 50, 0.5)
    ref_dt, ref_new = reference(nb, o, f, 50, 0.5)
    assert used == ref_dt and final == ref_new
    assert rechecked == [6, 7]


def test_planted_mistake_is_caught():
    """The comparison detects a wrong protocol: a guess that ignores the
    updated blocks before its first block, with no check launch behind it,
    gives wrong dts on some cases."""
    wrong = 0
    for seed in range(300):
        nb, o, f, skip, kstep = random_case(seed)
        ref_dt, _ = reference(nb, o, f, kstep, 0.5)
        used, _, _, _ = three_launches(nb, o, f, skip, kstep, 0.5, plant=True)
        wrong += used[1:] != ref_dt[1:]
    assert wrong > 0
