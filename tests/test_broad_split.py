"""The proximity pass's candidates from the first CCD pass's tree walk.

The device walks the tree once per step, over the swept boxes of the first
CCD pass, and takes the proximity pass's candidates from that walk
(k_traverse<DUAL>, k_split_seed).  That is exact when
  * every element's static box (the proximity pass's box: its points at x_old,
    inflated by 1e-6; dcollid.cpp:852-932) lies inside its swept box (x_old
    and x_old + dt * avgVel, same inflation), so a static overlap is always a
    swept overlap, and
  * two elements that share a vertex always overlap in their static boxes (the
    walk counts such rejected pairs without loading their static boxes).
Both are checked here in FP64 with the device's formulas on a small scene of
the config-4 kind, and the static candidate set is compared with the swept
candidate set filtered by static overlap (all pairs, brute force).
"""
import numpy as np

from collision_amd import scenes as S

EPS = 1e-6


def element_points(sc):
    ne = sc.nt + sc.nb
    idx = np.full((ne, 3), -1, np.int64)
    idx[:sc.nt] = sc.tris
    idx[sc.nt:, :2] = sc.bonds
    return idx


def boxes(pts, idx):
    """min/max over each element's points (bonds: two), inflated by EPS"""
    p = pts[np.where(idx >= 0, idx, idx[:, :1])]  # a bond repeats its first point
    return p.min(axis=1) - EPS, p.max(axis=1) + EPS


def overlap_pairs(lo, hi):
    n = lo.shape[0]
    out = set()
    for i in range(n - 1):
        m = np.all((lo[i] <= hi[i + 1:]) & (hi[i] >= lo[i + 1:]), axis=1)
        out.update((i, i + 1 + j) for j in np.nonzero(m)[0])
    return out


def test_static_boxes_inside_swept_boxes_and_shared_vertices_overlap():
    sc = S.combined_scene(n=20, nstrings=6, steps=2)
    rng = np.random.default_rng(3)
    x_old = sc.x.astype(np.float64)
    # velocities after impulses: displacements up to a fraction of an element
    av = sc.vel + rng.normal(0.0, 40.0, sc.vel.shape)
    dt = sc.dt
    idx = element_points(sc)
    # static boxes: triangles at x_old, bonds at Coords, which equals x_old in
    # the proximity pass
    slo, shi = boxes(x_old, idx)
    # swept boxes: x_old and x_old + dt * avgVel
    p0 = x_old[np.where(idx >= 0, idx, idx[:, :1])]
    p1 = p0 + dt * av[np.where(idx >= 0, idx, idx[:, :1])]
    wlo = np.minimum(p0, p1).min(axis=1) - EPS
    whi = np.maximum(p0, p1).max(axis=1) + EPS
    assert np.all(wlo <= slo) and np.all(whi >= shi)

    static = overlap_pairs(slo, shi)
    swept = overlap_pairs(wlo, whi)
    assert static <= swept
    filtered = {(i, j) for (i, j) in swept
                if np.all((slo[i] <= shi[j]) & (shi[i] >= slo[j]))}
    assert filtered == static

    # pairs with a shared vertex overlap in their static boxes
    ne = idx.shape[0]
    owners = {}
    for e in range(ne):
        for v in idx[e]:
            if v >= 0:
                owners.setdefault(int(v), []).append(e)
    shared = {(min(a, b), max(a, b)) for es in owners.values() for a in es for b in es if a != b}
    assert shared and shared <= static
