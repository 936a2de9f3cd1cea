"""GPU parity: the HIP collision step against the CPU oracle, bitwise.

Every test drives both through the reference's per-step loop (predict,
resolveCollision) on the same seeded scene and compares, after each step,
the per-pass candidate / narrow / hit counts, the set of hit element pairs
with their sub-test masks, the impact-zone membership, and the complete
vertex state (positions, velocities, average velocities, has_collsn), all
bit for bit.
"""
import numpy as np
import pytest

from collision_amd import scenes as S
from tests.parity import run_parity

pytestmark = pytest.mark.gpu


def test_config0_full_300_steps():
    """Config 0: 32x32 cloth onto a static 642-vertex icosphere, 300 steps.
    Contact starts near step 249 and includes CCD passes, an impact-zone
    escalation and strain limiting."""
    st = run_parity(S.config0(), steps=300, verbose=True)
    assert st["hit_pairs"] > 1000
    assert st["zone_steps"] >= 1


def test_config1_reduced_self_collision():
    """Config 1 at 64x64 (same generator, amplitude and velocity law)."""
    sc = S.self_collision_cloth(n=64, steps=12)
    st = run_parity(sc, steps=12, verbose=True)
    assert st["hit_pairs"] > 0


@pytest.mark.parametrize("full", [0, 1], ids=["incremental", "full_passes"])
def test_config4_reduced_cloth_strings_boxes(full):
    """Config 4 geometry at 96x96 cloth, 60 strings, friction 0.3; with the
    incremental later passes (default) and with every pass recomputed."""
    sc = S.combined_scene(n=96, nstrings=60, steps=60)
    st = run_parity(sc, steps=60, check_every=5, verbose=True, gpu_kw=dict(force_full_passes=full))
    assert st["hit_pairs"] > 0
    assert st["zone_steps"] > 0


@pytest.mark.parametrize("maker", [S.deck_ball_ball, S.deck_string_box, S.deck_string_string,
                                   S.deck_string_on_static_box])
def test_reference_deck_scenarios(maker):
    sc = maker()
    st = run_parity(sc, steps=min(sc.steps, 120), check_every=10)
    assert st["steps"] > 0
