"""ctypes binding of the CPU oracle (cs_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Loaded by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

ORDER_CANONICAL = 0
ORDER_REFERENCE_TREE = 1
MATH_DET = 0
MATH_LIBM = 1
PASS_NAMES = {0: "proximity", 1: "ccd", 2: "zone"}


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("thickness", C.c_double), ("round_eps", C.c_double),
                ("spring_k", C.c_double), ("mass", C.c_double), ("friction", C.c_double),
                ("restitution", C.c_double), ("dom_lo", C.c_double * 3),
                ("dom_hi", C.c_double * 3), ("max_ccd_passes", C.c_int32),
                ("max_zone_passes", C.c_int32), ("max_strain_sweeps", C.c_int32),
                ("order_mode", C.c_int32), ("math_mode", C.c_int32), ("log_pairs", C.c_int32),
                ("record_impulses", C.c_int32), ("mach_eps", C.c_double)]


class PassInfo(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_int32), ("candidates", C.c_int64),
                ("narrow_pairs", C.c_int64), ("hit_pairs", C.c_int64),
                ("hit_subtests", C.c_int64)]


class Report(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_passes", C.c_int32), ("n_ccd_passes", C.c_int32),
                ("n_zone_passes", C.c_int32), ("has_collision", C.c_int32),
                ("n_zones", C.c_int32), ("strain_sweeps", C.c_int32),
                ("strain_edges", C.c_int32), ("degenerate_edges", C.c_int32),
                ("pad", C.c_int32), ("candidates", C.c_int64), ("narrow_pairs", C.c_int64),
                ("hit_pairs", C.c_int64), ("hit_subtests", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


def build(force: bool = False) -> str:
    src = [os.path.join(HERE, f) for f in ("cs_oracle.c", "cs_oracle.h", "orc_detmath.h")]
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < max(os.path.getmtime(s) for s in src):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, i32, u8, i64 = C.c_double, C.c_int32, C.c_uint8, C.c_int64
        L.orc_default_params.argtypes = [P(Params)]
        L.orc_create.argtypes = [P(Params)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_params.argtypes = [C.c_void_p, P(Params)]
        L.orc_assemble.argtypes = [C.c_void_p, i32, i32, i32, P(i32), P(i32), P(d), P(i32), P(d), P(u8)]
        L.orc_set_state.argtypes = [C.c_void_p, P(d), P(d), P(u8)]
        L.orc_get_state.argtypes = [C.c_void_p, P(d), P(d), P(d), P(d), P(u8)]
        L.orc_predict.argtypes = [C.c_void_p]
        L.orc_resolve.argtypes = [C.c_void_p, P(Report)]
        L.orc_pass_count.argtypes = [C.c_void_p]
        L.orc_pass_count.restype = i32
        L.orc_pass_infos.argtypes = [C.c_void_p, P(PassInfo), i32]
        L.orc_pair_log_count.argtypes = [C.c_void_p]
        L.orc_pair_log_count.restype = i64
        L.orc_pair_log.argtypes = [C.c_void_p, P(i32), i64]
        L.orc_zone_labels.argtypes = [C.c_void_p, P(i32)]
        L.orc_set_bodies.argtypes = [C.c_void_p, i32, P(i32), P(d), P(d)]
        L.orc_get_bodies.argtypes = [C.c_void_p, P(d), P(d)]
        L.orc_set_avg_vel.argtypes = [C.c_void_p, P(d)]
        L.orc_last_impulses.argtypes = [C.c_void_p, P(d), P(i32)]
        L.orc_debug_pair.argtypes = [C.c_void_p, i32, i32, i32, i32]
        L.orc_debug_pair.restype = i32
        L.orc_debug_make_set.argtypes = [C.c_void_p]
        L.orc_is_coplanar.argtypes = [P(d), P(d), d, i32, P(d)]
        for fn in (L.orc_point_to_tri, L.orc_edge_to_edge):
            fn.argtypes = [P(Params), P(d), P(d), P(d), P(u8), d, P(d), P(d), P(i32)]
        L.orc_zone_rigid_update.argtypes = [P(Params), i32, P(d), P(d), P(u8), i32]
        for fn in (L.orc_math_acos, L.orc_math_cos, L.orc_math_sin, L.orc_math_pow_third):
            fn.argtypes = [d]
            fn.restype = d
        L.orc_math_eval.argtypes = [i32, P(d), P(d), i64]
        L.orc_libm_misrounded.argtypes = [i32]
        L.orc_libm_misrounded.restype = i64
        _lib = L
    return _lib


def _ptr(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def default_params(**kw) -> Params:
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("dom_lo", "dom_hi"):
            setattr(p, k, (C.c_double * 3)(*v))
        else:
            setattr(p, k, v)
    return p


def params_for_scene(scene, **kw) -> Params:
    base = dict(dt=scene.dt, thickness=scene.thickness, round_eps=scene.round_eps,
                spring_k=scene.spring_k, mass=scene.mass, friction=scene.friction,
                restitution=scene.restitution, dom_lo=scene.domain_lo, dom_hi=scene.domain_hi)
    base.update(kw)
    return default_params(**base)


class OracleSolver:
    """CPU oracle with the reference's driver loop: predict(); resolve()."""

    def __init__(self, scene, **kw):
        self.L = lib()
        self.scene = scene
        self.params = params_for_scene(scene, **kw)
        self.h = self.L.orc_create(C.byref(self.params))
        if not self.h:
            raise MemoryError("orc_create failed")
        self.nv = scene.nv
        tris = np.ascontiguousarray(scene.tris, np.int32)
        surf = np.ascontiguousarray(scene.tri_surf, np.int32)
        l0 = np.ascontiguousarray(scene.tri_len0, np.float64)
        bonds = np.ascontiguousarray(scene.bonds, np.int32)
        bl0 = np.ascontiguousarray(scene.bond_len0, np.float64)
        fixed = np.ascontiguousarray(scene.fixed, np.uint8)
        rc = self.L.orc_assemble(self.h, scene.nv, scene.nt, scene.nb, _ptr(tris, C.c_int32),
                                 _ptr(surf, C.c_int32), _ptr(l0, C.c_double),
                                 _ptr(bonds, C.c_int32), _ptr(bl0, C.c_double),
                                 _ptr(fixed, C.c_uint8))
        if rc:
            raise ValueError(f"orc_assemble failed: {rc}")
        self.set_state(scene.x, scene.vel)
        self.nbody = getattr(scene, "nbody", 0)
        if self.nbody:
            vb = np.ascontiguousarray(scene.vert_body, np.int32)
            bm = np.ascontiguousarray(scene.body_mass, np.float64)
            bc = np.ascontiguousarray(scene.body_com, np.float64)
            rc = self.L.orc_set_bodies(self.h, self.nbody, _ptr(vb, C.c_int32), _ptr(bm, C.c_double),
                                       _ptr(bc, C.c_double))
            if rc:
                raise ValueError(f"orc_set_bodies failed: {rc}")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, x, vel, has_collsn=None):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(vel, np.float64)
        hc = None if has_collsn is None else np.ascontiguousarray(has_collsn, np.uint8)
        self.L.orc_set_state(self.h, _ptr(x, C.c_double), _ptr(v, C.c_double), _ptr(hc, C.c_uint8))

    def get_state(self):
        n = self.nv
        x = np.empty((n, 3)); v = np.empty((n, 3)); av = np.empty((n, 3)); xo = np.empty((n, 3))
        hc = np.empty(n, np.uint8)
        self.L.orc_get_state(self.h, _ptr(x, C.c_double), _ptr(v, C.c_double),
                             _ptr(av, C.c_double), _ptr(xo, C.c_double), _ptr(hc, C.c_uint8))
        return dict(x=x, vel=v, avg_vel=av, x_old=xo, has_collsn=hc)

    def predict(self):
        self.L.orc_predict(self.h)

    def resolve(self) -> Report:
        rep = Report()
        self.L.orc_resolve(self.h, C.byref(rep))
        return rep

    def step(self) -> Report:
        self.predict()
        return self.resolve()

    def passes(self):
        n = self.L.orc_pass_count(self.h)
        arr = (PassInfo * max(n, 1))()
        self.L.orc_pass_infos(self.h, arr, n)
        return [dict(kind=PASS_NAMES[p.kind], index=p.index, candidates=p.candidates,
                     narrow_pairs=p.narrow_pairs, hit_pairs=p.hit_pairs,
                     hit_subtests=p.hit_subtests) for p in arr[:n]]

    def pair_log(self) -> np.ndarray:
        n = self.L.orc_pair_log_count(self.h)
        out = np.zeros((max(n, 0), 4), np.int32)
        if n:
            self.L.orc_pair_log(self.h, _ptr(out, C.c_int32), n)
        return out

    def zone_labels(self) -> np.ndarray:
        out = np.empty(self.nv, np.int32)
        self.L.orc_zone_labels(self.h, _ptr(out, C.c_int32))
        return out

    def bodies(self):
        """(center_of_mass, center_of_mass_velo) of the movable bodies"""
        com = np.zeros((max(self.nbody, 1), 3))
        vel = np.zeros((max(self.nbody, 1), 3))
        self.L.orc_get_bodies(self.h, _ptr(com, C.c_double), _ptr(vel, C.c_double))
        return com[:self.nbody], vel[:self.nbody]

    def last_impulses(self):
        imp, num = np.empty((self.nv, 3)), np.empty(self.nv, np.int32)
        if self.L.orc_last_impulses(self.h, _ptr(imp, C.c_double), _ptr(num, C.c_int32)):
            raise ValueError("record_impulses was not set")
        return imp, num

    # -- test hooks --
    def set_avg_vel(self, av):
        av = np.ascontiguousarray(av, np.float64)
        self.L.orc_set_avg_vel(self.h, _ptr(av, C.c_double))

    def debug_make_set(self):
        self.L.orc_debug_make_set(self.h)

    def debug_pair(self, A, B, ccd=1, zone_mode=0):
        """hit mask of one pair through the narrow phase (-1: error)"""
        return self.L.orc_debug_pair(self.h, A, B, ccd, zone_mode)


# ---- unit-level helpers ----
MATH_FUNCS = {"acos": 0, "cos": 1, "sin": 2, "pow_third": 3}


def libm_misrounded(reset=False):
    """libm-mode calls since the last reset whose libm result was not the
    correctly rounded one"""
    return lib().orc_libm_misrounded(1 if reset else 0)


def math_eval(name, x, libm=False):
    """vectorised deterministic-mode (or host libm) acos / cos / sin / pow(x, 1/3)"""
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    fn = MATH_FUNCS[name] + (10 if libm else 0)
    rc = lib().orc_math_eval(fn, _ptr(x, C.c_double), _ptr(y, C.c_double), x.size)
    if rc:
        raise ValueError(name)
    return y

def is_coplanar(x_old, avg_vel, dt, math_mode=MATH_DET):
    xo = np.ascontiguousarray(x_old, np.float64).reshape(4, 3)
    av = np.ascontiguousarray(avg_vel, np.float64).reshape(4, 3)
    roots = np.zeros(4)
    ok = lib().orc_is_coplanar(_ptr(xo, C.c_double), _ptr(av, C.c_double), dt, math_mode,
                               _ptr(roots, C.c_double))
    return bool(ok), roots


def _unit(fn, params, c, x_old, avg_vel, fixed, h):
    c = np.ascontiguousarray(c, np.float64).reshape(4, 3)
    xo = np.ascontiguousarray(x_old, np.float64).reshape(4, 3)
    av = np.ascontiguousarray(avg_vel, np.float64).reshape(4, 3)
    fx = np.ascontiguousarray(np.zeros(4) if fixed is None else fixed, np.uint8)
    imp = np.zeros((4, 3)); fric = np.zeros((4, 3)); cnt = np.zeros(4, np.int32)
    r = fn(C.byref(params), _ptr(c, C.c_double), _ptr(xo, C.c_double), _ptr(av, C.c_double),
           _ptr(fx, C.c_uint8), h, _ptr(imp, C.c_double), _ptr(fric, C.c_double),
           _ptr(cnt, C.c_int32))
    return r, imp, fric, cnt


def point_to_tri(params, c, x_old, avg_vel, fixed=None, h=None):
    h = params.thickness if h is None else h
    return _unit(lib().orc_point_to_tri, params, c, x_old, avg_vel, fixed, h)


def edge_to_edge(params, c, x_old, avg_vel, fixed=None, h=None):
    h = params.thickness if h is None else h
    return _unit(lib().orc_edge_to_edge, params, c, x_old, avg_vel, fixed, h)


def zone_rigid_update(params, x_old, avg_vel, fixed=None, lanes=0):
    xo = np.ascontiguousarray(x_old, np.float64)
    av = np.array(avg_vel, np.float64, copy=True, order="C")
    n = xo.shape[0]
    fx = np.ascontiguousarray(np.zeros(n) if fixed is None else fixed, np.uint8)
    rc = lib().orc_zone_rigid_update(C.byref(params), n, _ptr(xo, C.c_double),
                                     _ptr(av, C.c_double), _ptr(fx, C.c_uint8), lanes)
    return rc, av
