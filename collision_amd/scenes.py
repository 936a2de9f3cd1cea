"""Deterministic synthetic scenes for the collision step.

Every array is produced with IEEE-754 basic operations only (+, -, *, /,
sqrt, floor) on float64 and wrapping uint64 integer arithmetic, so a scene is
bit-identical on any host.  The transcendental functions the generators need
(log for Box-Muller, sin/cos for the noise angle and the height field) are
evaluated by the series in this module, not by a platform libm.

Mesh conventions (the reference library defines none; they are fixed here):
  * cloth n x n: vertex (i, j) -> id i*n + j at (x0 + i*s, y0 + j*s, z),
    s = side/(n-1); cell (i, j) gives triangles (v00, v10, v11) and
    (v00, v11, v01); rest lengths are the t=0 edge lengths.
  * icosphere: icosahedron plus L midpoint subdivisions projected to the
    radius; L=2 -> 162 vertices / 320 triangles, L=3 -> 642 / 1280.
  * box: 8 vertices / 12 triangles.
  * string: straight polyline of m bonds; rest length of every bond is the
    t=0 bond length.
Element order follows the reference's assembly order (surface triangles in
surface order, then the bonds of string curves; dcollid3d.cpp:25-43), and
vertex ids follow the same object order.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

U64 = np.uint64
_MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
SEED_BASE = 0xC0111D00

PI = 3.141592653589793
TWO_PI = 6.283185307179586
LN2 = 0.6931471805599453


# ----------------------------------------------------------------------------
# counter-based SplitMix64 stream
# ----------------------------------------------------------------------------
def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """z_k = mix(seed + (offset + k + 1) * GOLDEN), k = 0..count-1."""
    k = np.arange(offset + 1, offset + count + 1, dtype=U64)
    with np.errstate(over="ignore"):
        z = U64(seed & _MASK) + k * U64(GOLDEN)
        z = (z ^ (z >> U64(30))) * U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> U64(27))) * U64(0x94D049BB133111EB)
        z = z ^ (z >> U64(31))
    return z


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Uniform doubles in (0, 1]: ((z >> 11) + 1) * 2^-53 (exact)."""
    z = splitmix64(seed, count, offset)
    return ((z >> U64(11)) + U64(1)).astype(np.float64) * (2.0 ** -53)


# ----------------------------------------------------------------------------
# deterministic elementary functions (series, basic ops only)
# ----------------------------------------------------------------------------
def det_log(u: np.ndarray) -> np.ndarray:
    """Natural log for u > 0: u = m 2^e with m in [sqrt(.5), sqrt(2)),
    log(m) = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.172, series to s^29."""
    u = np.asarray(u, dtype=np.float64)
    m, e = np.frexp(u)                      # m in [0.5, 1)
    small = m < 0.7071067811865476
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e).astype(np.float64)
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    acc = np.full_like(s, 1.0 / 29.0)
    for k in range(27, 0, -2):
        acc = acc * s2 + 1.0 / k
    return 2.0 * s * acc + e * LN2


def _sin_poly(x: np.ndarray) -> np.ndarray:
    # Taylor series of sin on |x| <= pi/4, terms through x^23
    x2 = x * x
    acc = np.zeros_like(x)
    fact = [1.0]
    for n in range(1, 24):
        fact.append(fact[-1] * n)
    for k in range(23, 0, -2):
        sign = -1.0 if ((k - 1) // 2) % 2 else 1.0
        acc = acc * x2 + sign / fact[k]
    return acc * x


def _cos_poly(x: np.ndarray) -> np.ndarray:
    # Taylor series of cos on |x| <= pi/4, terms through x^22
    x2 = x * x
    acc = np.zeros_like(x)
    fact = [1.0]
    for n in range(1, 24):
        fact.append(fact[-1] * n)
    for k in range(22, -1, -2):
        sign = -1.0 if (k // 2) % 2 else 1.0
        acc = acc * x2 + sign / fact[k]
    return acc


def det_sincos_cycles(t: np.ndarray):
    """(sin(2 pi t), cos(2 pi t)) for t in [0, 1) given in cycles."""
    t = np.asarray(t, dtype=np.float64)
    q = np.floor(t * 4.0 + 0.5)            # nearest quarter turn
    r = t - q * 0.25                       # exact, |r| <= 1/8
    x = r * TWO_PI                         # |x| <= pi/4
    s, c = _sin_poly(x), _cos_poly(x)
    qi = q.astype(np.int64) & 3
    sin_t = np.select([qi == 0, qi == 1, qi == 2], [s, c, -s], -c)
    cos_t = np.select([qi == 0, qi == 1, qi == 2], [c, -s, -c], s)
    return sin_t, cos_t


def normal(seed: int, count: int, sigma: float, offset: int = 0) -> np.ndarray:
    """Box-Muller: sqrt(-2 ln u1) cos(2 pi u2) with u1, u2 from interleaved
    counters 2k, 2k+1 of the stream."""
    u = uniform01(seed, 2 * count, 2 * offset)
    u1, u2 = u[0::2], u[1::2]
    radius = np.sqrt(-2.0 * det_log(u1))
    _, c = det_sincos_cycles(u2 - np.floor(u2))
    return sigma * radius * c


# ----------------------------------------------------------------------------
# scene container
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class Scene:
    name: str
    x: np.ndarray            # (nv, 3) float64 positions at t=0
    vel: np.ndarray          # (nv, 3) float64 velocities
    fixed: np.ndarray        # (nv,) uint8, 1 = static rigid body point
    tris: np.ndarray         # (nt, 3) int32
    tri_surf: np.ndarray     # (nt,) int32 surface id
    tri_len0: np.ndarray     # (nt, 3) float64 rest lengths of edges (j, j+1)
    bonds: np.ndarray        # (nb, 2) int32
    bond_len0: np.ndarray    # (nb,) float64
    thickness: float
    dt: float
    steps: int
    friction: float = 0.0
    round_eps: float = 1e-6
    spring_k: float = 1000.0
    mass: float = 0.01
    restitution: float = 0.0
    domain_lo: tuple = (-10.0, -10.0, -10.0)
    domain_hi: tuple = (10.0, 10.0, 10.0)
    scene_of_vertex: Optional[np.ndarray] = None   # batched scenes only
    elem_scene: Optional[np.ndarray] = None        # batched scenes: scene of every element
    # movable rigid bodies (MOVABLE_BODY_BOUNDARY surfaces): body of every
    # vertex (-1: none), total mass and centre of mass of every body
    vert_body: Optional[np.ndarray] = None
    body_mass: Optional[np.ndarray] = None
    body_com: Optional[np.ndarray] = None
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def nbody(self):
        return 0 if self.body_mass is None else len(self.body_mass)

    @property
    def nv(self):
        return self.x.shape[0]

    @property
    def nt(self):
        return self.tris.shape[0]

    @property
    def nb(self):
        return self.bonds.shape[0]


class _Builder:
    def __init__(self):
        self.xs: List[np.ndarray] = []
        self.vs: List[np.ndarray] = []
        self.fx: List[np.ndarray] = []
        self.tris: List[np.ndarray] = []
        self.surf: List[np.ndarray] = []
        self.bonds: List[np.ndarray] = []
        self.body: List[np.ndarray] = []
        self.bmass: List[float] = []
        self.bcom: List[np.ndarray] = []
        self.nv = 0
        self.nsurf = 0

    def add_surface(self, x, v, tris, fixed, movable_mass=None):
        """A surface; static when `fixed`, a movable rigid body of total mass
        `movable_mass` (centre of mass = mean of its vertices) when given."""
        base = self.nv
        self.xs.append(np.asarray(x, np.float64))
        self.vs.append(np.asarray(v, np.float64))
        self.fx.append(np.full(len(x), 1 if fixed else 0, np.uint8))
        if movable_mass is not None:
            assert not fixed
            self.body.append(np.full(len(x), len(self.bmass), np.int32))
            self.bmass.append(float(movable_mass))
            self.bcom.append(np.asarray(x, np.float64).mean(axis=0))
        else:
            self.body.append(np.full(len(x), -1, np.int32))
        self.tris.append(np.asarray(tris, np.int32) + base)
        self.surf.append(np.full(len(tris), self.nsurf, np.int32))
        self.nsurf += 1
        self.nv += len(x)
        return base

    def add_string(self, x, v):
        base = self.nv
        m = len(x) - 1
        self.xs.append(np.asarray(x, np.float64))
        self.vs.append(np.asarray(v, np.float64))
        self.fx.append(np.zeros(len(x), np.uint8))
        self.body.append(np.full(len(x), -1, np.int32))
        k = np.arange(m, dtype=np.int32)
        self.bonds.append(np.stack([k + base, k + 1 + base], axis=1))
        self.nv += len(x)
        return base

    def build(self, name, **kw) -> Scene:
        x = np.concatenate(self.xs) if self.xs else np.zeros((0, 3))
        v = np.concatenate(self.vs) if self.vs else np.zeros((0, 3))
        fixed = np.concatenate(self.fx) if self.fx else np.zeros(0, np.uint8)
        tris = np.concatenate(self.tris) if self.tris else np.zeros((0, 3), np.int32)
        surf = np.concatenate(self.surf) if self.surf else np.zeros(0, np.int32)
        bonds = np.concatenate(self.bonds) if self.bonds else np.zeros((0, 2), np.int32)
        bodies = {}
        if self.bmass:
            bodies = dict(vert_body=np.ascontiguousarray(np.concatenate(self.body), np.int32),
                          body_mass=np.array(self.bmass, np.float64),
                          body_com=np.ascontiguousarray(np.array(self.bcom, np.float64)))
        return Scene(name=name, x=np.ascontiguousarray(x), vel=np.ascontiguousarray(v),
                     fixed=fixed, tris=np.ascontiguousarray(tris, dtype=np.int32),
                     tri_surf=surf, tri_len0=tri_rest_lengths(x, tris),
                     bonds=np.ascontiguousarray(bonds, dtype=np.int32),
                     bond_len0=bond_rest_lengths(x, bonds), **bodies, **kw)


def edge_length(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    d = a - b
    return np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])


def tri_rest_lengths(x, tris):
    tris = np.asarray(tris)
    if len(tris) == 0:
        return np.zeros((0, 3))
    out = np.empty((len(tris), 3))
    for j in range(3):
        out[:, j] = edge_length(x[tris[:, j]], x[tris[:, (j + 1) % 3]])
    return out


def bond_rest_lengths(x, bonds):
    bonds = np.asarray(bonds)
    if len(bonds) == 0:
        return np.zeros(0)
    return edge_length(x[bonds[:, 0]], x[bonds[:, 1]])


# ----------------------------------------------------------------------------
# primitive meshes
# ----------------------------------------------------------------------------
def cloth_grid(n: int, side: float, origin=(0.0, 0.0), z=0.0):
    s = side / (n - 1)
    i = np.repeat(np.arange(n, dtype=np.float64), n)
    j = np.tile(np.arange(n, dtype=np.float64), n)
    x = np.empty((n * n, 3))
    x[:, 0] = origin[0] + i * s
    x[:, 1] = origin[1] + j * s
    x[:, 2] = z
    ci = np.repeat(np.arange(n - 1), n - 1)
    cj = np.tile(np.arange(n - 1), n - 1)
    v00 = ci * n + cj
    v10 = (ci + 1) * n + cj
    v11 = (ci + 1) * n + cj + 1
    v01 = ci * n + cj + 1
    tris = np.empty((2 * (n - 1) * (n - 1), 3), np.int32)
    tris[0::2] = np.stack([v00, v10, v11], axis=1)
    tris[1::2] = np.stack([v00, v11, v01], axis=1)
    return x, tris


def icosphere(level: int, radius: float, center=(0.0, 0.0, 0.0)):
    phi = (1.0 + np.sqrt(5.0)) / 2.0
    base = np.array([[-1, phi, 0], [1, phi, 0], [-1, -phi, 0], [1, -phi, 0],
                     [0, -1, phi], [0, 1, phi], [0, -1, -phi], [0, 1, -phi],
                     [phi, 0, -1], [phi, 0, 1], [-phi, 0, -1], [-phi, 0, 1]],
                    dtype=np.float64)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [tuple(v) for v in base]

    def unit(v):
        v = np.asarray(v, np.float64)
        n = np.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
        return v / n

    verts = [unit(v) for v in verts]
    for _ in range(level):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = (verts[a] + verts[b]) * 0.5
                verts.append(unit(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    v = np.array(verts) * radius + np.asarray(center, np.float64)
    return v, np.array(faces, np.int32)


def box_mesh(center, edge):
    cx, cy, cz = center
    hx, hy, hz = edge[0] * 0.5, edge[1] * 0.5, edge[2] * 0.5
    v = np.array([[cx - hx, cy - hy, cz - hz], [cx + hx, cy - hy, cz - hz],
                  [cx + hx, cy + hy, cz - hz], [cx - hx, cy + hy, cz - hz],
                  [cx - hx, cy - hy, cz + hz], [cx + hx, cy - hy, cz + hz],
                  [cx + hx, cy + hy, cz + hz], [cx - hx, cy + hy, cz + hz]])
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7],
                  [0, 1, 5], [0, 5, 4], [1, 2, 6], [1, 6, 5],
                  [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int32)
    return v, f


def polyline(start, end, segments):
    start = np.asarray(start, np.float64)
    end = np.asarray(end, np.float64)
    k = np.arange(segments + 1, dtype=np.float64)[:, None]
    return start + k * ((end - start) / segments)


# ----------------------------------------------------------------------------
# the five workload configurations (+ reduced variants for parity tests)
# ----------------------------------------------------------------------------
def seed_for(config: int, sub: int = 0) -> int:
    return (SEED_BASE + config + (sub << 32)) & _MASK


def cloth_over_sphere(n=32, side=1.0, z=0.5, radius=0.25, level=3,
                      noise=1e-3, seed=None, thickness=1e-3, dt=1e-3, steps=300,
                      name="config0_cloth_sphere"):
    """One n x n cloth centred over a static icosphere at the origin;
    cloth velocity (0,0,-1) + N(0, noise) per component."""
    seed = seed_for(0) if seed is None else seed
    b = _Builder()
    x, t = cloth_grid(n, side, origin=(-side / 2, -side / 2), z=z)
    v = np.zeros_like(x)
    v[:, 2] = -1.0
    v += normal(seed, x.size, noise).reshape(-1, 3)
    b.add_surface(x, v, t, fixed=False)
    sx, st = icosphere(level, radius)
    b.add_surface(sx, np.zeros_like(sx), st, fixed=True)
    return b.build(name, thickness=thickness, dt=dt, steps=steps)


def config0():
    return cloth_over_sphere()


def self_collision_cloth(n=256, side=1.0, k=8, amp=0.004, znoise=2e-4, vsig=0.5,
                         vclip=2.0, seed=None, thickness=1e-3, dt=1e-3, steps=100,
                         name="config1_self_collision"):
    """Wavy n x n cloth with random velocities N(0, vsig) clipped to vclip."""
    seed = seed_for(1) if seed is None else seed
    x, t = cloth_grid(n, side, origin=(0.0, 0.0), z=0.0)
    tx = x[:, 0] * k
    ty = x[:, 1] * k
    sx, _ = det_sincos_cycles(tx - np.floor(tx))
    sy, _ = det_sincos_cycles(ty - np.floor(ty))
    x[:, 2] = amp * sx * sy + normal(seed, len(x), znoise)
    v = normal(seed ^ 0x5555, x.size, vsig).reshape(-1, 3)
    v = np.clip(v, -vclip, vclip)
    b = _Builder()
    b.add_surface(x, v, t, fixed=False)
    return b.build(name, thickness=thickness, dt=dt, steps=steps)


def config1(n=256):
    return self_collision_cloth(n=n)


def combined_scene(n=1024, side=2.0, nboxes=4, nstrings=2000, segs=64, noise=1e-3,
                   seed=None, thickness=5e-4, dt=5e-4, steps=100, friction=0.3,
                   name="config4_cloth_strings_boxes"):
    """Cloth n x n (side `side`, z = 0) falling at (0,0,-1)+N(0,noise) onto
    `nboxes` static boxes whose tops lie 5-20 mm below it; `nstrings` straight
    strings of `segs` bonds lie 2-10 mm above the cloth, falling at (0,0,-1).
    Placements scale with `side` so reduced variants keep the same geometry."""
    seed = seed_for(4) if seed is None else seed
    half = side / 2.0
    b = _Builder()
    x, t = cloth_grid(n, side, origin=(-half, -half), z=0.0)
    v = np.zeros_like(x)
    v[:, 2] = -1.0
    v += normal(seed, x.size, noise).reshape(-1, 3)
    b.add_surface(x, v, t, fixed=False)
    u = uniform01(seed ^ 0xB0B0, 6 * nboxes)
    for i in range(nboxes):
        ux, uy, ex, ey, ut, _ = u[6 * i:6 * i + 6]
        cx = (ux * 1.4 - 0.7) * half
        cy = (uy * 1.4 - 0.7) * half
        edx = (0.2 + 0.3 * ex) * half
        edy = (0.2 + 0.3 * ey) * half
        top = -(0.005 + 0.015 * ut)
        h = 0.3 * half
        bx, bt = box_mesh((cx, cy, top - h / 2), (edx, edy, h))
        b.add_surface(bx, np.zeros_like(bx), bt, fixed=True)
    us = uniform01(seed ^ 0x5757, 5 * nstrings)
    for s in range(nstrings):
        ux, uy, ua, ul, uz = us[5 * s:5 * s + 5]
        cx = (ux * 1.8 - 0.9) * half
        cy = (uy * 1.8 - 0.9) * half
        sa, ca = det_sincos_cycles(np.array([ua * 0.5]))
        length = (0.15 + 0.15 * ul) * half
        z = 0.002 + 0.008 * uz
        dx, dy = ca[0] * length * 0.5, sa[0] * length * 0.5
        p = polyline((cx - dx, cy - dy, z), (cx + dx, cy + dy, z), segs)
        vs = np.zeros_like(p)
        vs[:, 2] = -1.0
        b.add_string(p, vs)
    return b.build(name, thickness=thickness, dt=dt, steps=steps, friction=friction)


def config4(n=1024, nstrings=2000):
    return combined_scene(n=n, nstrings=nstrings)


def batched_scenes(count=512, nmin=16, nmax=128, seed=None, spacing=2.0,
                   thickness=1e-3, dt=1e-3, steps=200, name="config2_batched"):
    """`count` independent cloth-over-sphere scenes laid out on a square grid
    `spacing` apart, so their boxes can never overlap (no cross-scene pairs).
    Side vertex counts are log-uniform in [nmin, nmax]; sphere radius
    U(0.2, 0.35) with 162 or 642 vertices."""
    seed = seed_for(2) if seed is None else seed
    u = uniform01(seed, 3 * count)
    lo, hi = det_log(np.array([float(nmin)]))[0], det_log(np.array([float(nmax)]))[0]
    b = _Builder()
    cols = int(np.ceil(np.sqrt(count)))
    sizes, owner, tri_owner = [], [], []
    for s in range(count):
        un, ur, ul = u[3 * s:3 * s + 3]
        n = int(np.floor(_det_exp(lo + (hi - lo) * un) + 0.5))
        n = min(max(n, nmin), nmax)
        r = 0.2 + 0.15 * ur
        level = 2 if ul < 0.5 else 3
        ox, oy = (s % cols) * spacing, (s // cols) * spacing
        x, t = cloth_grid(n, 1.0, origin=(ox - 0.5, oy - 0.5), z=0.5)
        v = np.zeros_like(x)
        v[:, 2] = -1.0
        v += normal(seed ^ (0x1000 + s), x.size, 1e-3).reshape(-1, 3)
        b.add_surface(x, v, t, fixed=False)
        sx, st = icosphere(level, r, center=(ox, oy, 0.0))
        b.add_surface(sx, np.zeros_like(sx), st, fixed=True)
        owner += [s] * (len(x) + len(sx))
        tri_owner += [s] * (len(t) + len(st))
        sizes.append((n, level, r))
    sc = b.build(name, thickness=thickness, dt=dt, steps=steps)
    sc.scene_of_vertex = np.array(owner, np.int32)
    sc.elem_scene = np.array(tri_owner, np.int32)
    sc.meta["scenes"] = sizes
    return sc


def scene_parts(sc: Scene) -> List[Scene]:
    """The scenes of a batch (`elem_scene`, triangles only) as separate
    scenes, vertices and elements renumbered in their batch order."""
    assert sc.elem_scene is not None and sc.nb == 0
    out = []
    for s in range(int(sc.elem_scene.max()) + 1):
        tsel = np.flatnonzero(sc.elem_scene == s)
        vsel = np.flatnonzero(sc.scene_of_vertex == s)
        remap = np.full(sc.nv, -1, np.int64)
        remap[vsel] = np.arange(len(vsel))
        surf = sc.tri_surf[tsel]
        _, surf_local = np.unique(surf, return_inverse=True)
        out.append(Scene(name=f"{sc.name}_{s}", x=sc.x[vsel].copy(), vel=sc.vel[vsel].copy(),
                         fixed=sc.fixed[vsel].copy(),
                         tris=np.ascontiguousarray(remap[sc.tris[tsel]], np.int32),
                         tri_surf=np.ascontiguousarray(surf_local, np.int32),
                         tri_len0=sc.tri_len0[tsel].copy(), bonds=np.zeros((0, 2), np.int32),
                         bond_len0=np.zeros(0), thickness=sc.thickness, dt=sc.dt,
                         steps=sc.steps, friction=sc.friction, round_eps=sc.round_eps,
                         spring_k=sc.spring_k, mass=sc.mass, restitution=sc.restitution,
                         domain_lo=sc.domain_lo, domain_hi=sc.domain_hi))
    return out


def _det_exp(x: float) -> float:
    # exp for moderate x via 2^k * series(r), r = x - k ln2, |r| <= ln2/2
    k = np.floor(x / LN2 + 0.5)
    r = x - k * LN2
    acc, term = 1.0, 1.0
    for i in range(1, 25):
        term = term * r / i
        acc = acc + term
    return float(np.ldexp(acc, int(k)))


# ----------------------------------------------------------------------------
# small scenarios restating the reference's six decks (behaviour tests only)
# ----------------------------------------------------------------------------
def deck_ball_ball(level=2, dt=1e-3, steps=200):
    """Two deformable spheres moving toward each other (in-ball_ball)."""
    b = _Builder()
    for c, vv in (((0.3, 0.3, 0.3), -0.2), ((0.2, 0.2, 0.2), 0.2)):
        x, t = icosphere(level, 0.05, center=c)
        b.add_surface(x, np.full_like(x, vv), t, fixed=False)
    return b.build("deck_ball_ball", thickness=1e-3, dt=dt, steps=steps)


def deck_string_box(dt=1e-3, steps=200, segs=40):
    """A string under a falling deformable slab (in-string_box)."""
    b = _Builder()
    x, t = box_mesh((0.25, 0.25, 0.3), (0.15, 0.15, 0.03))
    v = np.zeros_like(x)
    v[:, 2] = -0.2
    b.add_surface(x, v, t, fixed=False)
    p = polyline((0.25, 0.01, 0.25), (0.25, 0.49, 0.25), segs)
    b.add_string(p, np.zeros_like(p))
    return b.build("deck_string_box", thickness=1e-3, dt=dt, steps=steps)


def deck_string_string(dt=1e-3, steps=150, segs=40):
    """Two crossing strings approaching each other (in-string_string)."""
    b = _Builder()
    p = polyline((0.25, 0.01, 0.24), (0.25, 0.49, 0.24), segs)
    v = np.zeros_like(p)
    v[:, 2] = 0.2
    b.add_string(p, v)
    q = polyline((0.01, 0.25, 0.25), (0.49, 0.25, 0.25), segs)
    w = np.zeros_like(q)
    w[:, 2] = -0.2
    b.add_string(q, w)
    return b.build("deck_string_string", thickness=1e-3, dt=dt, steps=steps)


def deck_string_on_static_box(dt=1e-3, steps=150, segs=40):
    """A string sliding down onto a static cuboid (in-string_slide with the
    body held static)."""
    b = _Builder()
    x, t = box_mesh((0.25, 0.25, 0.13), (0.2, 0.2, 0.02))
    b.add_surface(x, np.zeros_like(x), t, fixed=True)
    p = polyline((0.25, 0.1, 0.16), (0.25, 0.25, 0.145), segs)
    v = np.zeros_like(p)
    v[:, 1] = 0.1
    v[:, 2] = -0.1
    b.add_string(p, v)
    return b.build("deck_string_slide_static", thickness=1e-3, dt=dt, steps=steps,
                   friction=0.3)


def deck_ball_plane(level=2, dt=1e-3, steps=120):
    """A movable rigid ball rising into a falling deformable slab
    (in-ball_plane: sphere at (0.25, 0.25, 0.4), r 0.05, velocity (0, 0, 0.2),
    MOVABLE_BODY_BOUNDARY; cuboid at z 0.47, 0.15 x 0.15 x 0.01, velocity
    (0, 0, -0.2)); domain [0, 0.5] x [0, 0.5] x [0.25, 0.75]."""
    b = _Builder()
    x, t = icosphere(level, 0.05, center=(0.25, 0.25, 0.4))
    v = np.zeros_like(x)
    v[:, 2] = 0.2
    b.add_surface(x, v, t, fixed=False, movable_mass=0.5)
    bx, bt = box_mesh((0.25, 0.25, 0.47), (0.15, 0.15, 0.01))
    bv = np.zeros_like(bx)
    bv[:, 2] = -0.2
    b.add_surface(bx, bv, bt, fixed=False)
    return b.build("deck_ball_plane", thickness=1e-3, dt=dt, steps=steps,
                   domain_lo=(0.0, 0.0, 0.25), domain_hi=(0.5, 0.5, 0.75))


def deck_box_boundary(dt=1e-3, steps=150):
    """A movable rigid box (in-box_boundary: centre (0.25, 0.25, 0.21), edges
    0.1 x 0.1 x 0.2, velocity (0.4, 0, -0.4)) falling onto a static plate
    (contacts between rigid points only: impulses spread over the body) and
    sliding into a small deformable cloth lying on the plate (mixed
    contacts); domain [0, 1] x [0, 0.5] x [0, 0.5]."""
    b = _Builder()
    x, t = box_mesh((0.25, 0.25, 0.21), (0.1, 0.1, 0.2))
    v = np.tile([0.4, 0.0, -0.4], (len(x), 1))
    b.add_surface(x, v, t, fixed=False, movable_mass=2.0)
    px, pt = box_mesh((0.3, 0.25, 0.08), (0.4, 0.3, 0.04))
    b.add_surface(px, np.zeros_like(px), pt, fixed=True)
    cx, ct = cloth_grid(6, 0.06, origin=(0.33, 0.22), z=0.1025)
    cv = np.zeros_like(cx)
    cv[:, 2] = -0.05
    b.add_surface(cx, cv, ct, fixed=False)
    return b.build("deck_box_boundary", thickness=1e-3, dt=dt, steps=steps, friction=0.2,
                   domain_lo=(0.0, 0.0, 0.0), domain_hi=(1.0, 0.5, 0.5))


def deck_string_slide(dt=1e-3, steps=150, segs=40):
    """A string sliding down onto a movable cuboid (in-string_slide: curve
    from (0.25, 0.1, 0.3) to (0.25, 0.25, 0.155) with velocity (0, 0.1, -0.1);
    cuboid at (0.25, 0.25, 0.13), 0.2 x 0.2 x 0.02, at rest,
    MOVABLE_BODY_BOUNDARY); the string starts lower so contact comes within
    the steps run."""
    b = _Builder()
    x, t = box_mesh((0.25, 0.25, 0.13), (0.2, 0.2, 0.02))
    b.add_surface(x, np.zeros_like(x), t, fixed=False, movable_mass=1.0)
    p = polyline((0.25, 0.1, 0.16), (0.25, 0.25, 0.145), segs)
    v = np.zeros_like(p)
    v[:, 1] = 0.1
    v[:, 2] = -0.1
    b.add_string(p, v)
    return b.build("deck_string_slide", thickness=1e-3, dt=dt, steps=steps, friction=0.3,
                   domain_lo=(0.0, 0.0, 0.0), domain_hi=(0.5, 0.5, 0.5))


def domain_wall_strings(dt=1e-3, steps=40, segs=12):
    """Strings driven into a tight domain box, one of them starting outside
    it on one axis with an inward velocity, so that the domain-box friction of
    one visit can make the next visit of the same point touch a face."""
    b = _Builder()
    p = polyline((0.02, 0.1, 0.1), (0.02, 0.3, 0.1), segs)
    v = np.tile([-0.9, 0.3, 0.0], (len(p), 1))
    b.add_string(p, v)
    # x_old below the x face; the first visit touches the y face only, its
    # friction shortens the inward x move, and the second visit touches x
    q = polyline((-0.0001, 0.4999, 0.1), (-0.0001, 0.4999, 0.3), segs)
    w = np.tile([0.105, 0.5, -2.5], (len(q), 1))
    b.add_string(q, w)
    r = polyline((0.1, 0.1, 0.45), (0.3, 0.1, 0.45), segs)
    u = np.tile([0.1, -0.4, 1.2], (len(r), 1))
    b.add_string(r, u)
    return b.build("domain_wall_strings", thickness=1e-3, dt=dt, steps=steps, friction=0.4,
                   domain_lo=(0.0, 0.0, 0.0), domain_hi=(0.5, 0.5, 0.5))


def fan_disc_on_box(nfan=96, rings=3, vsig=2.0, dt=1e-3, steps=8, seed=7):
    """A flat disc meshed as a fan of `nfan` triangles around one centre
    vertex plus `rings` - 1 bands, with noisy velocities, falling onto a
    static box.  The centre vertex is shared by `nfan` elements, so its
    visits in a strain sweep span at least `nfan` levels."""
    b = _Builder()
    s, c = det_sincos_cycles(np.arange(nfan) / nfan)
    pts = [np.array([[0.0, 0.0, 0.01]])]
    for r in range(1, rings + 1):
        pts.append(np.stack([0.02 * r * c, 0.02 * r * s, np.full(nfan, 0.01)], 1))
    x = np.concatenate(pts)
    tris = []
    for i in range(nfan):
        tris.append((0, 1 + i, 1 + (i + 1) % nfan))
    for r in range(1, rings):
        o0, o1 = 1 + (r - 1) * nfan, 1 + r * nfan
        for i in range(nfan):
            j = (i + 1) % nfan
            tris.append((o0 + i, o1 + i, o1 + j))
            tris.append((o0 + i, o1 + j, o0 + j))
    v = normal(seed, x.size, vsig).reshape(-1, 3)
    v[:, 2] -= 1.0
    b.add_surface(x, v, np.array(tris, np.int32), fixed=False)
    bx, bt = box_mesh((0.0, 0.0, -0.1), (0.3, 0.3, 0.1))
    b.add_surface(bx, np.zeros_like(bx), bt, fixed=True)
    return b.build("fan_disc_on_box", thickness=1e-3, dt=dt, steps=steps)


CONFIGS = {
    "config0": config0,
    "config1": config1,
    "config2": batched_scenes,
    "config4": config4,
}
