"""Case model: the reference JSON schema, seeded generators, geometry."""
import json

import numpy as np

from hyperflow_amd import case as C


def test_json_round_trip():
    c = C.step_channel(16)
    c2 = C.Case.from_json(json.loads(json.dumps(c.to_json())))
    assert c2.to_json() == c.to_json()
    for a, b in zip(c.initial_fields(), c2.initial_fields()):
        assert np.array_equal(a, b)


def test_quadrant_seeds_are_deterministic_and_in_range():
    s1 = C.quadrant_states(np.random.default_rng(11))
    s2 = C.quadrant_states(np.random.default_rng(11))
    assert s1 == s2
    for q in s1.values():
        assert 0.1 <= q[0] <= 2.0 and 0.1 <= q[3] <= 2.0
        assert -1.0 <= q[1] <= 1.0 and -1.0 <= q[2] <= 1.0


def test_step_channel_internal_faces_are_consistent():
    c = C.step_channel(8)
    opposite = {"left": "right", "right": "left", "bottom": "top", "top": "bottom"}
    for b, blk in enumerate(c.blocks):
        for side, bc in blk.boundaries.items():
            if bc["type"] == "Internal":
                nb = c.blocks[bc["block"]]
                back = nb.boundaries[opposite[side]]
                assert back["type"] == "Internal" and back["block"] == b
    assert c.interior_cells == 8 * 64


def test_perturbed_fields_keep_density_positive():
    c = C.muscl_quadrants(2, 32)
    f = C.perturbed_fields(c, 2)[0]
    g = c.ghost
    assert (f[0, g:-g, g:-g] > 0).all()
    assert not np.array_equal(f, c.initial_fields()[0])


def test_batched_sizes():
    s = C.batched_sizes(512, 0)
    assert len(s) == 512
    assert all(64 <= a <= 512 and 64 <= b <= 512 for a, b in s)
