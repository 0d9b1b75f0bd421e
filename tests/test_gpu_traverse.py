"""The two-pass BVH traversal against the one-pass walk, on the 256 x 256 sphere drop
through its late, pair-dense substeps (most leaves at the 64-pair cap, the longest
walks near 900 visits): bodies that differ only in the first pass's visit budget
(CLOTHMI_TRAV_BUDGET: 0 = one pass; 160 = the default; 4 = nearly every leaf handed
to the wave-wide second pass, past the hand-over records' capacity the rest walking
on in the first pass) step the same drop in lockstep. Their pair lists (order
included), warnings and positions must be identical at every checked substep."""
import os

import numpy as np
import pytest

from licenta_amd import configs

pytestmark = pytest.mark.gpu


def _body(scene, budget):
    from licenta_amd.body import ClothBody
    old = os.environ.get("CLOTHMI_TRAV_BUDGET")
    os.environ["CLOTHMI_TRAV_BUDGET"] = str(budget)
    try:
        return ClothBody.from_scene(scene)
    finally:
        if old is None:
            del os.environ["CLOTHMI_TRAV_BUDGET"]
        else:
            os.environ["CLOTHMI_TRAV_BUDGET"] = old


@pytest.mark.timeout(300)
def test_two_pass_traversal_matches_one_pass():
    sc = configs.config_scene(1)
    budgets = (0, 160, 4)
    bodies = [_body(sc, bg) for bg in budgets]
    try:
        for b in bodies:
            b.set_state(sc["xyz"], np.zeros_like(sc["xyz"]))
        checks = list(range(100, 1201, 100)) + [1150, 1175, 1199]
        done = 0
        for stop in sorted(set(checks)):
            for b in bodies:
                b.fixed_update(stop - done)
            done = stop
            ref = bodies[0].pairs()
            print(f"[traverse] substep {stop}: {len(ref)} pairs", flush=True)
            for bg, b in zip(budgets[1:], bodies[1:]):
                got = b.pairs()
                assert len(got) == len(ref), (stop, bg)
                assert np.array_equal(got, ref), (stop, bg)
                assert b.warnings() == bodies[0].warnings(), (stop, bg)
                assert np.array_equal(b.positions(), bodies[0].positions()), (stop, bg)
        assert len(ref) > 1_000_000  # the dense regime was reached
    finally:
        for b in bodies:
            b.close()


def _batch(scenes, budget):
    from licenta_amd.body import ClothBody
    old = os.environ.get("CLOTHMI_TRAV_BUDGET")
    os.environ["CLOTHMI_TRAV_BUDGET"] = str(budget)
    try:
        return ClothBody.batch(scenes)
    finally:
        if old is None:
            del os.environ["CLOTHMI_TRAV_BUDGET"]
        else:
            os.environ["CLOTHMI_TRAV_BUDGET"] = old


@pytest.mark.timeout(300)
def test_two_pass_traversal_batch():
    """The same in a batch of 32 uneven cloths (per-cloth roots and caps, the
    smallest cloths' caps below the body's): budgets 0, 2 and 160 in lockstep over
    300 substeps."""
    scenes = configs.batched_scene(2, count=32)
    budgets = (0, 2, 160)
    bodies = [_batch(scenes, bg) for bg in budgets]
    try:
        x0 = np.concatenate([s["xyz"] for s in scenes])
        for b in bodies:
            b.set_state(x0, np.zeros_like(x0))
        for k in range(6):
            for b in bodies:
                b.fixed_update(50)
            ref = bodies[0].pairs()
            print(f"[traverse batch] substep {50 * (k + 1)}: {len(ref)} pairs", flush=True)
            for bg, b in zip(budgets[1:], bodies[1:]):
                assert np.array_equal(b.pairs(), ref), (k, bg)
                assert b.warnings() == bodies[0].warnings(), (k, bg)
                assert np.array_equal(b.positions(), bodies[0].positions()), (k, bg)
        assert len(ref) > 0
    finally:
        for b in bodies:
            b.close()
