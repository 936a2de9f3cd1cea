"""Python host API of the collision step, mirroring the reference's
CollisionSolver / CollisionSolver3d (collid.h:128-244) over the C ABI.

    solver = CollisionSolver3d()
    solver.setFabricThickness(h); solver.setTimeStepSize(dt); ...
    solver.assembleFromMesh(scene)          # assembleFromInterface
    solver.setState(x, vel)
    for each step:
        solver.predict()                    # the driver's x <- x_old + dt v
        solver.resolveCollision()           # dcollid.cpp:317-362

Setters are per instance (the reference's are class statics, dcollid.cpp:56-89).
Errors raise CollisionError carrying the C status instead of exiting the
process as the reference's clean_up(ERROR) does.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MeshView, Params, PassInfo, StepReport


class CollisionError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_lib.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def default_params(**kw) -> Params:
    p = Params()
    _lib.lib().cs_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("dom_lo", "dom_hi"):
            setattr(p, k, (C.c_double * 3)(*v))
        else:
            setattr(p, k, v)
    return p


class CollisionSolver3d:
    def __init__(self, params: Params | None = None, **kw):
        self.L = _lib.lib()
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        rc = self.L.cs_create(C.byref(h), C.byref(self.params))
        if rc:
            raise CollisionError(rc, self.L.cs_last_error(None).decode())
        self.h = h
        self.nv = 0
        self._keep = None
        self.last_report = None

    # -- lifecycle --
    def close(self):
        if getattr(self, "h", None):
            self.L.cs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, allow_warn=False):
        if rc and not (allow_warn and rc == _lib.CS_WARN_ZONE_CAP):
            raise CollisionError(rc, self.L.cs_last_error(self.h).decode())
        return rc

    def _push(self):
        self._check(self.L.cs_set_params(self.h, C.byref(self.params)))

    # -- the reference's static setters / getters (dcollid.cpp:56-89) --
    def setRoundingTolerance(self, v): self.params.round_eps = v; self._push()
    def getRoundingTolerance(self): return self.params.round_eps
    def setFabricThickness(self, v): self.params.thickness = v; self._push()
    def getFabricThickness(self): return self.params.thickness
    def setTimeStepSize(self, v): self.params.dt = v; self._push()
    def getTimeStepSize(self): return self.params.dt
    def setSpringConstant(self, v): self.params.spring_k = v; self._push()
    def getSpringConstant(self): return self.params.spring_k
    def setFrictionConstant(self, v): self.params.friction = v; self._push()
    def getFrictionConstant(self): return self.params.friction
    def setPointMass(self, v): self.params.mass = v; self._push()
    def getPointMass(self): return self.params.mass
    def setRestitutionCoef(self, v): self.params.restitution = v; self._push()
    # per-pass timing points (not in the reference): 2 phases and kernels, 1 kernels, 0 none
    def setPassTiming(self, level): self.params.pass_timing = level; self._push()
    def getRestitutionCoef(self): return self.params.restitution

    def setDomainBoundary(self, L, U):
        lo = np.ascontiguousarray(L, np.float64)
        hi = np.ascontiguousarray(U, np.float64)
        self._check(self.L.cs_set_domain_boundary(self.h, _p(lo, C.c_double), _p(hi, C.c_double)))
        self.params.dom_lo = (C.c_double * 3)(*lo)
        self.params.dom_hi = (C.c_double * 3)(*hi)

    def getDomainBoundary(self, dir, side):
        return (self.params.dom_lo if side == 0 else self.params.dom_hi)[dir]

    # -- topology and state --
    def assemble(self, nv, tris, tri_surf, tri_len0, bonds, bond_len0, is_fixed,
                 vert_body=None, body_mass=None, body_com=None, elem_scene=None):
        arrs = dict(tris=np.ascontiguousarray(tris, np.int32).reshape(-1, 3),
                    tri_surf=np.ascontiguousarray(tri_surf, np.int32),
                    tri_len0=np.ascontiguousarray(tri_len0, np.float64).reshape(-1, 3),
                    bonds=np.ascontiguousarray(bonds, np.int32).reshape(-1, 2),
                    bond_len0=np.ascontiguousarray(bond_len0, np.float64),
                    is_fixed=np.ascontiguousarray(is_fixed, np.uint8))
        nbody = 0 if body_mass is None else len(body_mass)
        if nbody:
            arrs.update(vert_body=np.ascontiguousarray(vert_body, np.int32),
                        body_mass=np.ascontiguousarray(body_mass, np.float64),
                        body_com=np.ascontiguousarray(body_com, np.float64).reshape(-1, 3))
        nscene = 0
        if elem_scene is not None:
            arrs["elem_scene"] = np.ascontiguousarray(elem_scene, np.int32)
            nscene = int(arrs["elem_scene"].max()) + 1 if arrs["elem_scene"].size else 0
        m = MeshView(nv, arrs["tris"].shape[0], arrs["bonds"].shape[0],
                     _p(arrs["tris"], C.c_int32), _p(arrs["tri_surf"], C.c_int32),
                     _p(arrs["tri_len0"], C.c_double), _p(arrs["bonds"], C.c_int32),
                     _p(arrs["bond_len0"], C.c_double), _p(arrs["is_fixed"], C.c_uint8),
                     nbody, _p(arrs.get("vert_body"), C.c_int32),
                     _p(arrs.get("body_mass"), C.c_double), _p(arrs.get("body_com"), C.c_double),
                     nscene, _p(arrs.get("elem_scene"), C.c_int32))
        self._check(self.L.cs_assemble(self.h, C.byref(m)))
        self.nv = nv
        self.nbody = nbody
        self.nscene = nscene

    def assembleFromMesh(self, scene):
        """assembleFromInterface over a flat scene (collision_amd.scenes.Scene)."""
        self.assemble(scene.nv, scene.tris, scene.tri_surf, scene.tri_len0, scene.bonds,
                      scene.bond_len0, scene.fixed, getattr(scene, "vert_body", None),
                      getattr(scene, "body_mass", None), getattr(scene, "body_com", None),
                      getattr(scene, "elem_scene", None))

    def setState(self, x, vel, has_collsn=None):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(vel, np.float64)
        hc = None if has_collsn is None else np.ascontiguousarray(has_collsn, np.uint8)
        self._check(self.L.cs_set_state(self.h, _p(x, C.c_double), _p(v, C.c_double),
                                        _p(hc, C.c_uint8)))

    def getState(self):
        n = self.nv
        out = dict(x=np.empty((n, 3)), vel=np.empty((n, 3)), avg_vel=np.empty((n, 3)),
                   x_old=np.empty((n, 3)), has_collsn=np.empty(n, np.uint8))
        self._check(self.L.cs_get_state(self.h, _p(out["x"], C.c_double),
                                        _p(out["vel"], C.c_double),
                                        _p(out["avg_vel"], C.c_double),
                                        _p(out["x_old"], C.c_double),
                                        _p(out["has_collsn"], C.c_uint8)))
        return out

    def recordOriginPosition(self):
        self._check(self.L.cs_record_origin_position(self.h))

    # -- step --
    def predict(self):
        self._check(self.L.cs_predict(self.h))

    def resolveCollision(self) -> StepReport:
        rep = StepReport()
        self._check(self.L.cs_resolve(self.h, C.byref(rep)), allow_warn=True)
        self.last_report = rep
        return rep

    def step(self) -> StepReport:
        self.predict()
        return self.resolveCollision()

    # -- launch shapes (no counterpart in the reference) --
    def tuning(self) -> dict:
        t = _lib.Tuning()
        self._check(self.L.cs_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def setTuning(self, traverse_blocks_per_cu, narrow_threads, query_grid):
        t = _lib.Tuning(traverse_blocks_per_cu, narrow_threads, query_grid, 0)
        self._check(self.L.cs_set_tuning(self.h, C.byref(t)))

    # -- batched scenes (no counterpart in the reference) --
    def resolveBatch(self):
        """cs_resolve_batch: (batch report, list of per-scene report dicts)"""
        rep = StepReport()
        per = (_lib.SceneReport * self.nscene)()
        self._check(self.L.cs_resolve_batch(self.h, C.byref(rep), per, self.nscene),
                    allow_warn=True)
        self.last_report = rep
        return rep, [r.as_dict() for r in per]

    def stepBatch(self):
        self.predict()
        return self.resolveBatch()

    def hasCollision(self) -> bool:
        return bool(self.L.cs_has_collision(self.h))

    def synchronize(self):
        self._check(self.L.cs_synchronize(self.h))

    # -- parity taps --
    def passes(self):
        n = self.L.cs_pass_count(self.h)
        arr = (PassInfo * max(n, 1))()
        self._check(self.L.cs_pass_infos(self.h, arr, n))
        return [dict(kind=_lib.PASS_NAMES[p.kind], index=p.index, candidates=p.candidates,
                     narrow_pairs=p.narrow_pairs, hit_pairs=p.hit_pairs,
                     hit_subtests=p.hit_subtests, tested_pairs=p.tested_pairs,
                     incremental=p.incremental, ms_broad=p.ms_broad, ms_narrow=p.ms_narrow,
                     ms_response=p.ms_response, ms_traverse_kernel=p.ms_traverse_kernel,
                     ms_narrow_kernel=p.ms_narrow_kernel) for p in arr[:n]]

    def pairLog(self) -> np.ndarray:
        n = self.L.cs_pair_log_count(self.h)
        out = np.zeros((max(n, 0), 4), np.int32)
        if n:
            self._check(self.L.cs_pair_log(self.h, _p(out, C.c_int32), n))
        return out

    def zoneLabels(self) -> np.ndarray:
        out = np.empty(self.nv, np.int32)
        self._check(self.L.cs_zone_labels(self.h, _p(out, C.c_int32)))
        return out

    def lastImpulses(self):
        """summed collsnImpulse and collsn_num of every vertex in the last pass"""
        imp, num = np.empty((self.nv, 3)), np.empty(self.nv, np.int32)
        self._check(self.L.cs_last_impulses(self.h, _p(imp, C.c_double), _p(num, C.c_int32)))
        return imp, num

    def writeVtk(self, path):
        """vtkplotVectorSurface (vtk.cpp:14-93) of the current state"""
        self._check(self.L.cs_write_vtk(self.h, str(path).encode()))

    def save(self, path):
        self._check(self.L.cs_save(self.h, str(path).encode()))

    def load(self, path):
        self._check(self.L.cs_load(self.h, str(path).encode()))
        p = Params()
        self._check(self.L.cs_get_params(self.h, C.byref(p)))
        self.params = p

    def getBodies(self):
        """(center_of_mass, center_of_mass_velo) of the movable bodies"""
        nb = getattr(self, "nbody", 0)
        com, vel = np.zeros((max(nb, 1), 3)), np.zeros((max(nb, 1), 3))
        self._check(self.L.cs_get_bodies(self.h, _p(com, C.c_double), _p(vel, C.c_double)))
        return com[:nb], vel[:nb]

    # ---- learned time-of-impact proposer (no counterpart in the reference) ----
    def setLearnedWeights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != _lib.MLP_PARAMS:
            raise ValueError(f"expected {_lib.MLP_PARAMS} parameters, got {w.size}")
        self._check(self.L.cs_learned_set_weights(self.h, _p(w, C.c_float)))

    def learnedWeights(self) -> np.ndarray:
        w = np.empty(_lib.MLP_PARAMS, np.float32)
        self._check(self.L.cs_learned_get_weights(self.h, _p(w, C.c_float)))
        return w

    def learnedInit(self, seed: int = 0):
        self._check(self.L.cs_learned_init(self.h, seed))

    def learnedInfer(self, x) -> np.ndarray:
        """(logit, raw seed) of each row of x [n][24] on the device"""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 24)
        out = np.empty((x.shape[0], 2), np.float32)
        self._check(self.L.cs_learned_infer(self.h, _p(x, C.c_float), _p(out, C.c_float),
                                            x.shape[0]))
        return out

    def learnedGrad(self, x, exists, toi):
        """(mean loss, gradient) of one batch on the device"""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 24)
        e = np.ascontiguousarray(exists, np.float32)
        t = np.ascontiguousarray(toi, np.float32)
        g = np.empty(_lib.MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._check(self.L.cs_learned_grad(self.h, _p(x, C.c_float), _p(e, C.c_float),
                                           _p(t, C.c_float), x.shape[0], _p(g, C.c_float),
                                           C.byref(loss)))
        return float(loss.value), g

    def learnedCollect(self, capacity: int, stride: int = 1):
        self._check(self.L.cs_learned_collect(self.h, capacity, stride))

    def learnedDataset(self):
        n = self.L.cs_learned_dataset_size(self.h)
        if n < 0:
            raise CollisionError(_lib.CS_ERR_HIP, "dataset size")
        x = np.empty((n, 24), np.float32)
        e, t = np.empty(n, np.float32), np.empty(n, np.float32)
        self._check(self.L.cs_learned_dataset(self.h, _p(x, C.c_float), _p(e, C.c_float),
                                              _p(t, C.c_float), n))
        return x, e, t

    def learnedTrain(self, epochs: int, batch: int = 8192, lr: float = 1e-3, seed: int = 0):
        loss = np.empty(max(epochs, 1), np.float32)
        self._check(self.L.cs_learned_train(self.h, epochs, batch, lr, seed, _p(loss, C.c_float)))
        return loss[:epochs]

    def learnedStats(self) -> dict:
        st = _lib.LearnedStats()
        self._check(self.L.cs_learned_get_stats(self.h, C.byref(st)))
        return st.as_dict()


def math_eval(fn: str, x, device: int = 0) -> np.ndarray:
    """The step's own acos / cos / sin / pow(x, 1/3), evaluated on the device."""
    codes = {"acos": 0, "cos": 1, "sin": 2, "pow_third": 3}
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    rc = _lib.lib().cs_math_eval(device, codes[fn], _p(x, C.c_double), _p(y, C.c_double), x.size)
    if rc:
        raise CollisionError(rc, "cs_math_eval failed")
    return y


def solver_for_scene(scene, **kw) -> CollisionSolver3d:
    base = dict(dt=scene.dt, thickness=scene.thickness, round_eps=scene.round_eps,
                spring_k=scene.spring_k, mass=scene.mass, friction=scene.friction,
                restitution=scene.restitution, dom_lo=scene.domain_lo, dom_hi=scene.domain_hi)
    base.update(kw)
    s = CollisionSolver3d(**base)
    s.assembleFromMesh(scene)
    s.setState(scene.x, scene.vel)
    return s
