"""Deterministic scene generators: sizes of the five workloads, bit-stable
arrays (hashes committed under tests/golden/), and the series-based
elementary functions against numpy."""
import hashlib
import json
import os

import numpy as np

from collision_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "scene_hashes.json")


def digest(sc):
    h = hashlib.sha256()
    for a in (sc.x, sc.vel, sc.fixed, sc.tris, sc.tri_surf, sc.tri_len0, sc.bonds, sc.bond_len0):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_workload_sizes():
    c0 = S.config0()
    assert (c0.nv, c0.nt, c0.nb) == (1666, 3202, 0)
    c1 = S.config1()
    assert (c1.nv, c1.nt, c1.nb) == (65536, 130050, 0)
    c4 = S.config4()
    assert (c4.nv, c4.nt, c4.nb) == (1178608, 2093106, 128000)
    assert S.icosphere(2, 1.0)[0].shape[0] == 162 and S.icosphere(3, 1.0)[0].shape[0] == 642


def test_rest_lengths_are_initial_edge_lengths():
    sc = S.config0()
    e = np.linalg.norm(sc.x[sc.tris[:, 0]] - sc.x[sc.tris[:, 1]], axis=1)
    np.testing.assert_allclose(sc.tri_len0[:, 0], e, rtol=1e-15)


def test_scene_arrays_are_bit_stable():
    golden = json.load(open(GOLDEN))
    for name, maker in (("config0", S.config0), ("config1_n64", lambda: S.config1(n=64)),
                        ("config4_n96", lambda: S.combined_scene(n=96, nstrings=60)),
                        ("config2_16", lambda: S.batched_scenes(count=16))):
        assert digest(maker()) == golden[name], name


def test_series_functions_match_numpy():
    u = S.uniform01(9, 50000)
    np.testing.assert_allclose(S.det_log(u), np.log(u), rtol=2e-16 * 4, atol=1e-300)
    t = u - np.floor(u)
    s, c = S.det_sincos_cycles(t)
    np.testing.assert_allclose(s, np.sin(2 * np.pi * t), atol=2e-15)
    np.testing.assert_allclose(c, np.cos(2 * np.pi * t), atol=2e-15)
    n = S.normal(3, 400000, 1.0)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1) < 0.01


def test_fan_disc_has_one_high_valence_vertex():
    """The fan disc used by the one-sweep-in-flight strain case: nfan triangles
    around vertex 0, (rings - 1) bands of 2 * nfan triangles, a static box, and
    the same arrays on every call."""
    a, b = S.fan_disc_on_box(), S.fan_disc_on_box()
    assert digest(a) == digest(b)
    nfan, rings = 96, 3
    movable = ~a.fixed.astype(bool)[a.tris].any(axis=1)
    assert movable.sum() == nfan + 2 * nfan * (rings - 1)
    valence = np.bincount(a.tris[movable].ravel(), minlength=a.nv)
    assert valence[0] == nfan and np.sort(valence)[-2] <= 6
    # the disc lies above the box, so the first step has no contact
    box_top = a.x[a.fixed.astype(bool), 2].max()
    assert a.x[~a.fixed.astype(bool), 2].min() > box_top
