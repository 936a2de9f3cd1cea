"""Pins the oracle's narrow-phase and response restatement with checks the
mathematics fixes, chosen so that a dropped term, a flipped sign, a swapped
operand or a wrong branch fails one of them:

  isCoplanar  (dcollid3d.cpp:371-482): det[x_i + t v_i] built from matrices
              with prescribed eigen-structure, so the roots are known exactly
              (three-root trigonometric branch, one-root Cardano branch,
              quadratic and linear fallbacks), plus random cubics against their
              exact (rational-coefficient) roots.
  PointToTri / EdgeToEdge (778-922, 643-776): textbook distances and
              barycentric / segment parameters, rejection cases.
  impulses    (925-1300): conservation of linear momentum, direction and
              barycentric proportionality of the impulses, and the
              separating velocity a static obstacle leaves behind.
  zone rigid  (dcollid3d.cpp:70-200): translation and rotation are rigid
              motions the update must reproduce; distances and the mean
              velocity are preserved.
"""
import math

import numpy as np
import pytest

MACH_EPS = np.finfo(float).eps


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def coplanar(O, x_rel, v_rel, dt=1e-3, math_mode=0):
    """points 1..3 given relative to point 0 at the origin with zero velocity"""
    xo = np.vstack([np.zeros(3), np.asarray(x_rel, float)])
    av = np.vstack([np.zeros(3), np.asarray(v_rel, float)])
    return O.is_coplanar(xo, av, dt, math_mode)


@pytest.mark.parametrize("math_mode", [0, 1])
def test_three_real_roots_trig_branch(O, math_mode):
    t = [2e-4, 5e-4, 8e-4]
    # X + t I with X = -diag(t): det = (t - t1)(t - t2)(t - t3)
    ok, roots = coplanar(O, -np.diag(t), np.eye(3), math_mode=math_mode)
    assert ok
    np.testing.assert_allclose(roots[:3], np.array(t) - MACH_EPS, rtol=1e-12, atol=1e-18)
    assert roots[3] == 1e-3


def test_three_real_roots_with_general_velocity_matrix(O):
    """det(X + tV) = det(V) det(V^-1 X + t I): roots = -eig(V^-1 X)."""
    rng = np.random.default_rng(7)
    for _ in range(50):
        V = rng.normal(size=(3, 3))
        t = np.sort(rng.uniform(1e-4, 9e-4, 3))
        X = -V @ np.diag(t)   # V^-1 X = -diag(t)
        ok, roots = coplanar(O, X, V)
        assert ok
        got = np.sort(roots[:3][roots[:3] >= 0])
        np.testing.assert_allclose(got, t - MACH_EPS, rtol=1e-7, atol=1e-13)


def test_single_real_root_cardano_branch(O):
    r, p = 4e-4, 1e-3
    # X + tI = [[t-r,0,0],[0,t,p],[0,-p,t]] -> (t - r)(t^2 + p^2)
    X = np.array([[-r, 0, 0], [0, 0, p], [0, -p, 0]])
    ok, roots = coplanar(O, X, np.eye(3))
    assert ok
    valid = roots[:3][roots[:3] >= 0]
    np.testing.assert_allclose(valid, [r - MACH_EPS], rtol=1e-12)


def test_quadratic_fallback(O):
    t1, t2, c = 3e-4, 7e-4, 0.5
    # v3 = 0: det(X + tV) = c (t - t1)(t - t2), leading coefficient det(V) = 0
    X = np.array([[-t1, 0, 0], [0, -t2, 0], [0, 0, c]])
    V = np.diag([1.0, 1.0, 0.0])
    ok, roots = coplanar(O, X, V)
    assert ok
    got = np.sort(roots[:3][roots[:3] >= 0])
    np.testing.assert_allclose(got, [t1 - MACH_EPS, t2 - MACH_EPS], rtol=1e-12)


def test_linear_fallback_point_falling_onto_static_triangle(O):
    gap, speed = 1e-4, 1.0
    xo = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.3, gap]], float)
    av = np.zeros((4, 3))
    av[3, 2] = -speed
    ok, roots = O.is_coplanar(xo, av, 1e-3, 0)
    assert ok
    valid = roots[:3][roots[:3] >= 0]
    np.testing.assert_allclose(valid, [gap / speed - MACH_EPS], rtol=1e-13)


def test_no_motion_no_root_and_root_at_zero_rejected(O):
    xo = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.3, 1e-4]], float)
    ok, _ = O.is_coplanar(xo, np.zeros((4, 3)), 1e-3, 0)
    assert not ok
    # already coplanar at t = 0 and separating: the root 0 is shifted below 0
    xo[3, 2] = 0.0
    av = np.zeros((4, 3))
    av[3, 2] = 1.0
    ok, roots = O.is_coplanar(xo, av, 1e-3, 0)
    assert not ok


def _exact_cubic(X, V):
    """Coefficients (t^3..t^0) of det(X + tV), exactly, from the double inputs."""
    from fractions import Fraction as F
    Xf = [[F(float(a)) for a in r] for r in X]
    Vf = [[F(float(a)) for a in r] for r in V]

    def polymul(p, q):
        out = [F(0)] * (len(p) + len(q) - 1)
        for i, a in enumerate(p):
            for j, b in enumerate(q):
                out[i + j] += a * b
        return out

    # entries as polynomials (constant, linear) in t; Leibniz expansion
    M = [[[Xf[i][j], Vf[i][j]] for j in range(3)] for i in range(3)]
    total = [F(0)] * 4
    for perm, sign in (((0, 1, 2), 1), ((1, 2, 0), 1), ((2, 0, 1), 1),
                       ((0, 2, 1), -1), ((2, 1, 0), -1), ((1, 0, 2), -1)):
        term = polymul(polymul(M[0][perm[0]], M[1][perm[1]]), M[2][perm[2]])
        for k in range(4):
            total[k] += sign * term[k]
    return total[::-1]


def test_random_cubics_match_exact_roots(O):
    """Random det(X + tV): the oracle's roots in (0, dt] against the exact
    cubic's roots (rational coefficients, high-precision Newton), within a
    tolerance scaled by each root's condition number."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(300):
        X = rng.normal(size=(3, 3)) * 1e-3
        V = rng.normal(size=(3, 3))
        c = _exact_cubic(X, V)
        cf = [float(a) for a in c]
        roots = []
        for r0 in np.roots(cf):
            if abs(r0.imag) > 1e-9 * max(1e-6, abs(r0)):
                continue
            t = Decimal(float(r0.real))
            cd = [Decimal(a.numerator) / Decimal(a.denominator) for a in c]
            for _ in range(60):
                p = ((cd[0] * t + cd[1]) * t + cd[2]) * t + cd[3]
                dp = (3 * cd[0] * t + 2 * cd[1]) * t + cd[2]
                if dp == 0:
                    break
                t = t - p / dp
            roots.append(float(t))
        roots = np.sort(np.array(roots))
        if len(roots) > 1 and np.min(np.diff(roots)) < 1e-6:
            continue   # near-double roots: ill-conditioned by construction
        # the closed form subtracts A/3 = -(r1 + r2 + r3)/3, so its absolute
        # error scales with the largest root magnitude; add the root's own
        # condition number for the coefficient rounding
        rmax = max(abs(r) for r in np.roots(cf))
        scale = lambda t: sum(abs(a) * abs(t) ** (3 - i) for i, a in enumerate(cf))
        dpdt = lambda t: abs(3 * cf[0] * t * t + 2 * cf[1] * t + cf[2])
        inr = [t for t in roots if 1e-9 < t - MACH_EPS < 1e-3 - 1e-9]
        ok, got = coplanar(O, X, V)
        got = np.sort(got[:3][got[:3] > 0])
        assert ok == (len(inr) > 0)
        assert len(got) == len(inr)
        for g, t in zip(got, inr):
            tol = 64 * MACH_EPS * (rmax + scale(t) / dpdt(t)) + 4 * MACH_EPS
            assert abs(g - (t - MACH_EPS)) <= tol, (g, t, tol)
        checked += 1
    assert checked > 250


# ---------------------------------------------------------------------------
def params(O, **kw):
    base = dict(dt=1e-3, thickness=1e-3, round_eps=1e-6, spring_k=1000.0, mass=0.01,
                friction=0.0)
    base.update(kw)
    return O.default_params(**base)


TRI = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)


def test_point_to_tri_hit_geometry_and_momentum(O):
    P = params(O)
    h = P.thickness
    c = np.vstack([TRI, [0.2, 0.3, 0.5 * h]])
    av = np.zeros((4, 3))
    av[3, 2] = -0.05                      # point approaching the triangle
    hit, imp, fric, cnt = O.point_to_tri(P, c, c, av)
    assert hit == 1
    assert list(cnt) == [1, 1, 1, 1]
    # linear momentum of the impulse is zero (weights sum to one)
    assert np.abs(imp.sum(axis=0)).max() < 1e-15 * np.abs(imp).max() * 10
    # all along the normal (z); the point is pushed up, the triangle down
    assert np.abs(imp[:, :2]).max() == 0.0
    assert imp[3, 2] > 0 and np.all(imp[:3, 2] < 0)
    # barycentric split: P = w0 p0 + w1 p1 + w2 p2 with (w0, w1, w2) = (0.5, 0.2, 0.3)
    np.testing.assert_allclose(imp[:3, 2] / imp[:3, 2].sum(), [0.5, 0.2, 0.3], rtol=1e-12)
    assert np.all(fric == 0)              # zero friction coefficient


def test_point_to_tri_rejections(O):
    P = params(O)
    h = P.thickness
    av = np.zeros((4, 3))
    far = np.vstack([TRI, [0.2, 0.3, 2 * h]])
    assert O.point_to_tri(P, far, far, av)[0] == 0
    outside = np.vstack([TRI, [0.8, 0.8, 0.5 * h]])       # w2 < 0
    assert O.point_to_tri(P, outside, outside, av)[0] == 0
    degenerate = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0.5, 0, 0.5 * h]], float)
    assert O.point_to_tri(P, degenerate, degenerate, av)[0] == 0


def test_static_obstacle_leaves_the_repulsion_velocity(O):
    """Point approaching a static triangle: the inelastic part removes the whole
    normal approach velocity and the repulsion spring leaves the point moving
    away at min(dt k d / m, 0.1 d / dt - vn), d = thickness - distance."""
    P = params(O)
    h, dt, k, m = P.thickness, P.dt, P.spring_k, P.mass
    for gap, vn in ((0.5 * h, -0.05), (0.9 * h, -2.0), (0.2 * h, -1e-4)):
        c = np.vstack([TRI, [0.2, 0.3, gap]])
        av = np.zeros((4, 3))
        av[3, 2] = vn
        hit, imp, fric, cnt = O.point_to_tri(P, c, c, av, fixed=[1, 1, 1, 0])
        assert hit == 1 and list(cnt) == [0, 0, 0, 1]
        assert np.all(imp[:3] == 0)
        d = h - gap
        v_after = vn + imp[3, 2] / cnt[3]
        np.testing.assert_allclose(v_after, min(dt * k * d / m, 0.1 * d / dt - vn), rtol=1e-12)


def test_edge_to_edge_hit_and_momentum(O):
    P = params(O)
    h = P.thickness
    z = 0.4 * h
    c = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, z], [0, 1, z]], float)
    av = np.zeros((4, 3))
    av[2:, 2] = -0.1                        # upper edge moving down
    hit, imp, fric, cnt = O.edge_to_edge(P, c, c, av)
    assert hit == 1 and list(cnt) == [1, 1, 1, 1]
    assert np.abs(imp.sum(axis=0)).max() < 1e-15
    # midpoints (a = b = 1/2): equal shares, lower edge pushed down, upper up
    np.testing.assert_allclose(imp[0], imp[1])
    np.testing.assert_allclose(imp[2], imp[3])
    assert imp[0, 2] < 0 < imp[2, 2]
    assert np.abs(imp[:, :2]).max() == 0.0


def test_edge_to_edge_rejections_and_clamping(O):
    P = params(O)
    h = P.thickness
    av = np.zeros((4, 3))
    parallel = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0.5 * h], [1, 0, 0.5 * h]], float)
    assert O.edge_to_edge(P, parallel, parallel, av)[0] == 0
    # infinite lines cross 0.5h apart, but the segments stop 0.1 short:
    # clamped parameters give a distance well above h
    apart = np.array([[0.1, 0, 0], [1, 0, 0], [0, 0.1, 0.5 * h], [0, 1, 0.5 * h]], float)
    assert O.edge_to_edge(P, apart, apart, av)[0] == 0
    # clamped to the segment ends (a = 0, b = 0) at distance 0.5h: a hit
    touch = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0.5 * h], [0, 1, 0.5 * h]], float)
    hit, imp, _, _ = O.edge_to_edge(P, touch, touch, av)
    assert hit == 1


def test_friction_opposes_tangential_motion(O):
    P = params(O, friction=0.3)
    h = P.thickness
    c = np.vstack([TRI, [0.2, 0.3, 0.5 * h]])
    av = np.zeros((4, 3))
    av[3] = [0.2, 0.0, -0.05]              # sliding in +x while approaching
    hit, imp, fric, cnt = O.point_to_tri(P, c, c, av, fixed=[1, 1, 1, 0])
    assert hit == 1
    assert fric[3, 0] < 0 and abs(fric[3, 1]) < 1e-18 and abs(fric[3, 2]) < 1e-18
    assert abs(fric[3, 0]) <= 0.2 + 1e-15   # Coulomb clamp: never reverses the slide


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lanes", [0, 64])
def test_zone_translation_is_preserved(O, lanes):
    P = params(O)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(150, 3)) * 0.01
    v = np.tile([0.3, -0.2, 0.1], (150, 1))
    rc, av = O.zone_rigid_update(P, x, v, lanes=lanes)
    assert rc == 0
    np.testing.assert_allclose(av, v, rtol=0, atol=1e-13)


@pytest.mark.parametrize("lanes", [0, 64])
def test_zone_rigid_rotation_and_projection(O, lanes):
    P = params(O)
    rng = np.random.default_rng(6)
    n = 130
    x = rng.normal(size=(n, 3)) * 0.1   # det(I) well above the 1e-10 cut
    cm = x.mean(axis=0)
    w = np.array([3.0, -1.0, 2.0])
    v = np.cross(w, x - cm) + np.array([0.1, 0.0, -0.2])
    rc, av = O.zone_rigid_update(P, x, v, lanes=lanes)
    assert rc == 0
    # the update is an exact rigid motion over dt: distances are preserved
    xn = x + P.dt * av
    d_old = np.linalg.norm(x[:, None] - x[None], axis=-1)
    d_new = np.linalg.norm(xn[:, None] - xn[None], axis=-1)
    np.testing.assert_allclose(d_new, d_old, rtol=1e-9, atol=1e-15)
    # mean velocity preserved; a rigid input field is reproduced to O(dt |w|)
    np.testing.assert_allclose(av.mean(axis=0), v.mean(axis=0), atol=1e-13)
    assert np.abs(av - v).max() < 2 * P.dt * np.linalg.norm(w) * np.abs(v).max()
    # a non-rigid field is projected onto a rigid one (distances still kept)
    v2 = v + rng.normal(size=v.shape) * 0.05
    rc, av2 = O.zone_rigid_update(P, x, v2, lanes=lanes)
    xn2 = x + P.dt * av2
    d2 = np.linalg.norm(xn2[:, None] - xn2[None], axis=-1)
    np.testing.assert_allclose(d2, d_old, rtol=1e-9, atol=1e-15)


def test_zone_small_inertia_drops_rotation(O):
    """det(I) < 1e-10 sets omega to zero (dcollid3d.cpp:133-134): a
    millimetre-sized zone moves as a pure translation at its mean velocity."""
    P = params(O)
    rng = np.random.default_rng(9)
    x = rng.normal(size=(40, 3)) * 0.001
    v = np.cross([3.0, -1.0, 2.0], x - x.mean(axis=0)) + 0.1
    rc, av = O.zone_rigid_update(P, x, v)
    np.testing.assert_allclose(av, np.tile(v.mean(axis=0), (40, 1)), atol=1e-13)


def test_zone_static_members_are_not_moved(O):
    P = params(O)
    x = np.array([[0, 0, 0], [1e-3, 0, 0], [0, 1e-3, 0], [0, 0, 1e-3]], float)
    v = np.array([[0, 0, 0], [0, 0, -1], [0, 0, -1], [0, 0, -1]], float)
    rc, av = O.zone_rigid_update(P, x, v, fixed=[1, 0, 0, 0])
    assert rc == 0
    assert np.all(av[0] == 0)


def test_zone_orders_agree(O):
    P = params(O)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(300, 3)) * 0.01
    v = rng.normal(size=(300, 3))
    _, a0 = O.zone_rigid_update(P, x, v, lanes=0)
    _, a64 = O.zone_rigid_update(P, x, v, lanes=64)
    np.testing.assert_allclose(a0, a64, rtol=1e-11, atol=1e-13)
