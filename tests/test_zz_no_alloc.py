"""cs_zone_labels_noalloc (staging in a fixed device array, no allocation)
against cs_zone_labels, through the C ABI, on two calls in a row, the second
on a smaller mesh."""
import ctypes as C

import numpy as np
import pytest

from collision_amd import _lib
from collision_amd import scenes as S
from collision_amd.solver import solver_for_scene

pytestmark = pytest.mark.gpu


def _labels(solver, fn):
    out = np.full(solver.nv, -7, np.int32)
    rc = fn(solver.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
    assert rc == 0, solver.L.cs_last_error(solver.h)
    return out


def test_zone_labels_noalloc_matches_zone_labels():
    L = _lib.lib()
    big = S.config0()
    big.x = big.x + big.vel * (big.dt * 236)
    a = solver_for_scene(big)
    for _ in range(64):                      # until a step with impact zones
        if a.step().n_zone_passes:
            break
    small = S.deck_ball_plane()              # 170 vertices: a body zone
    b = solver_for_scene(small)
    for _ in range(3):
        b.step()
    try:
        for s in (a, b):                     # second call: the smaller mesh
            ref = _labels(s, L.cs_zone_labels)
            got = _labels(s, L.cs_zone_labels_noalloc)
            assert np.array_equal(ref, got)
            assert (ref != np.arange(s.nv)).any()     # some zone holds two or more points
    finally:
        a.close()
        b.close()
