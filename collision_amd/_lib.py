"""ctypes binding of libcollision_amd.so (the C ABI in include/collision_amd.h).

Loading fails loudly if the library has not been built: there is no CPU
fallback for the collision step.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libcollision_amd.so")
HEADER = os.path.join(ROOT, "include", "collision_amd.h")

CS_OK = 0
CS_ERR_ARG = 1
CS_ERR_NONFINITE = 2
CS_ERR_UNSUPPORTED = 3
CS_ERR_HIP = 4
CS_ERR_NOMEM = 5
CS_ERR_NO_DEVICE = 6
CS_ERR_INTERNAL = 7
CS_WARN_ZONE_CAP = 16

STATUS_NAMES = {0: "CS_OK", 1: "CS_ERR_ARG", 2: "CS_ERR_NONFINITE", 3: "CS_ERR_UNSUPPORTED",
                4: "CS_ERR_HIP", 5: "CS_ERR_NOMEM", 6: "CS_ERR_NO_DEVICE", 7: "CS_ERR_INTERNAL",
                16: "CS_WARN_ZONE_CAP"}
PASS_NAMES = {0: "proximity", 1: "ccd", 2: "zone"}


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("thickness", C.c_double), ("round_eps", C.c_double),
                ("spring_k", C.c_double), ("mass", C.c_double), ("friction", C.c_double),
                ("restitution", C.c_double), ("dom_lo", C.c_double * 3),
                ("dom_hi", C.c_double * 3), ("max_ccd_passes", C.c_int32),
                ("max_zone_passes", C.c_int32), ("max_strain_sweeps", C.c_int32),
                ("log_pairs", C.c_int32), ("device", C.c_int32),
                ("force_full_passes", C.c_int32), ("pass_timing", C.c_int32),
                ("record_impulses", C.c_int32), ("precision", C.c_int32),
                ("learned", C.c_int32), ("precision_stats", C.c_int32), ("graph_step", C.c_int32),
                ("autotune", C.c_int32), ("pad2", C.c_int32)]


class MeshView(C.Structure):
    _fields_ = [("nv", C.c_int32), ("nt", C.c_int32), ("nb", C.c_int32),
                ("tris", C.POINTER(C.c_int32)), ("tri_surf", C.POINTER(C.c_int32)),
                ("tri_len0", C.POINTER(C.c_double)), ("bonds", C.POINTER(C.c_int32)),
                ("bond_len0", C.POINTER(C.c_double)), ("is_fixed", C.POINTER(C.c_uint8)),
                ("nbody", C.c_int32), ("vert_body", C.POINTER(C.c_int32)),
                ("body_mass", C.POINTER(C.c_double)), ("body_com", C.POINTER(C.c_double)),
                ("nscene", C.c_int32), ("elem_scene", C.POINTER(C.c_int32))]


class PassInfo(C.Structure):
    _fields_ = [("kind", C.c_int32), ("index", C.c_int32), ("candidates", C.c_int64),
                ("narrow_pairs", C.c_int64), ("hit_pairs", C.c_int64),
                ("hit_subtests", C.c_int64), ("tested_pairs", C.c_int64),
                ("incremental", C.c_int32), ("pad", C.c_int32), ("ms_broad", C.c_double),
                ("ms_narrow", C.c_double), ("ms_response", C.c_double),
                ("ms_traverse_kernel", C.c_double), ("ms_narrow_kernel", C.c_double)]


class StepReport(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_passes", C.c_int32), ("n_ccd_passes", C.c_int32),
                ("n_zone_passes", C.c_int32), ("has_collision", C.c_int32),
                ("n_zones", C.c_int32), ("strain_sweeps", C.c_int32),
                ("strain_edges", C.c_int32), ("degenerate_edges", C.c_int32),
                ("strain_levels", C.c_int32), ("candidates", C.c_int64), ("narrow_pairs", C.c_int64),
                ("hit_pairs", C.c_int64), ("hit_subtests", C.c_int64),
                ("ms_total", C.c_double), ("ms_broad", C.c_double), ("ms_narrow", C.c_double),
                ("ms_response", C.c_double), ("ms_zone", C.c_double), ("ms_strain", C.c_double),
                ("strain_slots", C.c_int64), ("ms_strain_kernel", C.c_double),
                ("host_syncs", C.c_int64), ("mx_screened", C.c_int64), ("mx_rechecked", C.c_int64),
                ("mx_err_dist", C.c_double), ("mx_err_param", C.c_double),
                ("mx_bound_ratio", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class Tuning(C.Structure):
    _fields_ = [("traverse_blocks_per_cu", C.c_int32), ("narrow_threads", C.c_int32),
                ("query_grid", C.c_int32), ("tuned", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SceneReport(C.Structure):
    _fields_ = [("n_passes", C.c_int32), ("n_ccd_passes", C.c_int32),
                ("n_zone_passes", C.c_int32), ("has_collision", C.c_int32),
                ("zone_cap", C.c_int32), ("pad", C.c_int32), ("hit_pairs", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class LearnedStats(C.Structure):
    _fields_ = [("scored", C.c_int64), ("exists", C.c_int64), ("predicted_exists", C.c_int64),
                ("agree", C.c_int64), ("newton_roots", C.c_int64), ("newton_iters", C.c_int64),
                ("newton_iters_mid", C.c_int64), ("max_root_err", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


MLP_PARAMS = 19970


def declared_symbols(header: str = HEADER):
    """Function names declared in the public header."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cs_[a-z_0-9]+)\s*\(", text)))


_lib = None


def lib(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with collision_amd.build.build() "
                           "(there is no CPU fallback for the collision step)")
    L = C.CDLL(path)
    P = C.POINTER
    vp, i32, i64, d, u8 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint8
    sig = {
        "cs_default_params": ([P(Params)], None),
        "cs_create": ([P(vp), P(Params)], C.c_int),
        "cs_destroy": ([vp], C.c_int),
        "cs_set_params": ([vp, P(Params)], C.c_int),
        "cs_get_params": ([vp, P(Params)], C.c_int),
        "cs_set_domain_boundary": ([vp, P(d), P(d)], C.c_int),
        "cs_assemble": ([vp, P(MeshView)], C.c_int),
        "cs_set_state": ([vp, P(d), P(d), P(u8)], C.c_int),
        "cs_get_state": ([vp, P(d), P(d), P(d), P(d), P(u8)], C.c_int),
        "cs_record_origin_position": ([vp], C.c_int),
        "cs_predict": ([vp], C.c_int),
        "cs_resolve": ([vp, P(StepReport)], C.c_int),
        "cs_step": ([vp, P(StepReport)], C.c_int),
        "cs_resolve_batch": ([vp, P(StepReport), P(SceneReport), i32], C.c_int),
        "cs_step_batch": ([vp, P(StepReport), P(SceneReport), i32], C.c_int),
        "cs_has_collision": ([vp], C.c_int),
        "cs_synchronize": ([vp], C.c_int),
        "cs_pass_count": ([vp], i32),
        "cs_pass_infos": ([vp, P(PassInfo), i32], C.c_int),
        "cs_pair_log_count": ([vp], i64),
        "cs_pair_log": ([vp, P(i32), i64], C.c_int),
        "cs_zone_labels": ([vp, P(i32)], C.c_int),
        "cs_zone_labels_noalloc": ([vp, P(i32)], C.c_int),
        "cs_get_bodies": ([vp, P(d), P(d)], C.c_int),
        "cs_math_eval": ([i32, i32, P(d), P(d), i64], C.c_int),
        "cs_last_impulses": ([vp, P(d), P(i32)], C.c_int),
        "cs_write_vtk": ([vp, C.c_char_p], C.c_int),
        "cs_save": ([vp, C.c_char_p], C.c_int),
        "cs_load": ([vp, C.c_char_p], C.c_int),
        "cs_learned_set_weights": ([vp, P(C.c_float)], C.c_int),
        "cs_learned_get_weights": ([vp, P(C.c_float)], C.c_int),
        "cs_learned_init": ([vp, C.c_uint64], C.c_int),
        "cs_learned_infer": ([vp, P(C.c_float), P(C.c_float), i64], C.c_int),
        "cs_learned_grad": ([vp, P(C.c_float), P(C.c_float), P(C.c_float), i64, P(C.c_float),
                             P(C.c_float)], C.c_int),
        "cs_learned_collect": ([vp, i64, i32], C.c_int),
        "cs_learned_dataset_size": ([vp], i64),
        "cs_learned_dataset": ([vp, P(C.c_float), P(C.c_float), P(C.c_float), i64], C.c_int),
        "cs_learned_train": ([vp, i32, i32, C.c_float, C.c_uint64, P(C.c_float)], C.c_int),
        "cs_learned_get_stats": ([vp, P(LearnedStats)], C.c_int),
        "cs_get_tuning": ([vp, P(Tuning)], C.c_int),
        "cs_set_tuning": ([vp, P(Tuning)], C.c_int),
        "cs_last_error": ([vp], C.c_char_p),
        "cs_version": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L
