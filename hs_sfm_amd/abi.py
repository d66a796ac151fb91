"""ctypes mirror of include/hsba.h.

Builds `hsba_problem_desc` records from `synthetic.Scene` arrays and loads the
HIP library `hs_sfm_amd/lib/libhsba.so`.  Loading fails loudly when the
library has not been built: there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional

import numpy as np

from .synthetic import Scene

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HSBA_LIB", os.path.join(HERE, "lib", "libhsba.so"))

HSBA_OK = 0
ERRORS = {
    -1: "HSBA_ERR_GENERIC", -2: "HSBA_ERR_INVALID", -3: "HSBA_ERR_NOT_SPD",
    -4: "HSBA_ERR_NAN", -5: "HSBA_ERR_OOM", -6: "HSBA_ERR_DEVICE",
    -7: "HSBA_ERR_ADMISSION", -8: "HSBA_ERR_UNSUPPORTED", -9: "HSBA_ERR_NOT_READY",
    -10: "HSBA_ERR_TOO_LARGE",
}
PRECISION_FP64 = 0
PRECISION_MIXED = 1
SOLVER_AUTO, SOLVER_DENSE, SOLVER_PCG = 0, 1, 2
LOSS_TRIVIAL, LOSS_HUBER = 0, 1
PAIRS_TUNED, PAIRS_VALU, PAIRS_MFMA, PAIRS_MFMA_WAVE = 0, 1, 2, 3
COV_AUTO, COV_SCHUR, COV_DENSE = 0, 1, 2
COV_DENSE_MAX = 4096

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class hsba_constraint(C.Structure):
    _fields_ = [("id", C.c_int64), ("mask", C.c_uint32), ("reserved_", C.c_uint32)]


class hsba_problem_desc(C.Structure):
    _fields_ = [
        ("num_points", C.c_int64), ("num_images", C.c_int64),
        ("num_cameras", C.c_int64), ("num_keys", C.c_int64),
        ("feature_maps", _ip), ("image_camera_map", _ip),
        ("intrinsic_computations_mask", C.c_uint32), ("intrinsic_free_mask", C.c_uint32),
        ("intrinsic_defaults", _dp),
        ("fix_mask", C.c_uint32), ("reserved_", C.c_uint32),
        ("fix_points", _dp), ("fix_images", _dp), ("fix_cameras", _dp),
        ("num_point_constraints", C.c_int64), ("point_constraints", C.POINTER(hsba_constraint)),
        ("num_image_constraints", C.c_int64), ("image_constraints", C.POINTER(hsba_constraint)),
        ("num_camera_constraints", C.c_int64), ("camera_constraints", C.POINTER(hsba_constraint)),
    ]


class hsba_structure_sizes(C.Structure):
    _fields_ = [("num_keys", C.c_int64), ("num_points", C.c_int64), ("num_point_image", C.c_int64),
                ("num_point_camera", C.c_int64), ("num_image_camera", C.c_int64),
                ("num_pairs", C.c_int64), ("num_incidences", C.c_int64)]


class hsba_ycov_inverse(C.Structure):
    _fields_ = [("key_blocks", _dp), ("constraint_weights", _dp),
                ("point_constraint_blocks", _dp)]


class hsba_options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("solver", C.c_int32), ("pcg_rtol", C.c_double),
                ("pcg_max_iterations", C.c_int32), ("point_group_lanes", C.c_int32),
                ("loss", C.c_int32), ("pair_kernel", C.c_int32), ("loss_scale", C.c_double)]


class hsba_tune_entry(C.Structure):
    _fields_ = [("precision", C.c_int32), ("camera_block", C.c_int32), ("bucket", C.c_int32),
                ("variant", C.c_int32), ("ms", C.c_double * 3), ("pairs", C.c_int64)]


class hsba_covariance_info(C.Structure):
    _fields_ = [("method", C.c_int32), ("rank", C.c_int32), ("sweeps", C.c_int32), ("reserved_", C.c_int32),
                ("smallest_kept", C.c_double)]


class hsba_lm_options(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("reserved_", C.c_int32), ("tau", C.c_double),
                ("eps1", C.c_double), ("eps2", C.c_double)]


class hsba_lm_stats(C.Structure):
    _fields_ = [("status", C.c_int32), ("stop_reason", C.c_int32), ("iterations", C.c_int32),
                ("accepted", C.c_int32), ("initial_cost", C.c_double), ("final_cost", C.c_double),
                ("final_mu", C.c_double), ("pcg_iterations", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class hsba_stage_times(C.Structure):
    _fields_ = [("linearize_ms", C.c_double), ("image_reduce_ms", C.c_double),
                ("point_factor_ms", C.c_double), ("pairs_ms", C.c_double),
                ("assemble_ms", C.c_double), ("cholesky_ms", C.c_double),
                ("backsub_ms", C.c_double), ("step_cost_ms", C.c_double),
                ("iteration_ms", C.c_double), ("iterations", C.c_int32), ("reserved_", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved_"}


class hsba_server_options(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active_jobs", C.c_int32),
                ("precision", C.c_int32), ("lm", hsba_lm_options), ("streams", C.c_int32),
                ("max_batch_jobs", C.c_int32), ("loss", C.c_int32), ("reserved2_", C.c_int32),
                ("loss_scale", C.c_double)]


class hsba_naive_desc(C.Structure):
    _fields_ = [("num_cameras", C.c_int64), ("num_points", C.c_int64), ("num_features", C.c_int64),
                ("feature_maps", C.POINTER(C.c_int32))]


class hsba_gcp_desc(C.Structure):
    _fields_ = [("num_cameras", C.c_int64), ("num_points", C.c_int64), ("num_features", C.c_int64),
                ("feature_maps", C.POINTER(C.c_int32)), ("num_gcps", C.c_int64)]


class hsba_sfm_files(C.Structure):
    _fields_ = [("tracks", C.c_char_p), ("keysets", C.POINTER(C.c_char_p)), ("num_keysets", C.c_int64),
                ("intrinsics", C.c_char_p), ("extrinsics", C.c_char_p), ("points", C.c_char_p),
                ("image_intrinsic_map", C.c_char_p), ("image_extrinsic_map", C.c_char_p),
                ("track_point_map", C.c_char_p)]


class hsba_scene_view(C.Structure):
    _fields_ = [("desc", hsba_problem_desc), ("x0", _dp), ("y_obs", _dp), ("x_size", C.c_int64),
                ("y_size", C.c_int64)]


class hsba_job_result(C.Structure):
    _fields_ = [("job_id", C.c_int64), ("status", C.c_int32), ("reserved_", C.c_int32),
                ("submit_time_s", C.c_double), ("admit_time_s", C.c_double),
                ("finish_time_s", C.c_double), ("stats", hsba_lm_stats)]


def dptr(a: Optional[np.ndarray]):
    if a is None:
        return C.cast(None, _dp)
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous required"
    return a.ctypes.data_as(_dp)


def iptr(a: Optional[np.ndarray]):
    if a is None:
        return C.cast(None, _ip)
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "int32 C-contiguous required"
    return a.ctypes.data_as(_ip)


class Desc:
    """Owns the numpy buffers a `hsba_problem_desc` points into."""

    def __init__(self, scene: Scene):
        self.scene = scene
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            b = np.ascontiguousarray(a, dtype=dt)
            keep.append(b)
            return b

        fm = arr(scene.feature_maps, np.int32)
        icm = arr(scene.image_camera_map, np.int32)
        d = hsba_problem_desc()
        d.num_points, d.num_images = scene.num_points, scene.num_images
        d.num_cameras, d.num_keys = scene.num_cameras, scene.num_keys
        d.feature_maps = fm.ctypes.data_as(_ip)
        d.image_camera_map = icm.ctypes.data_as(_ip)
        d.intrinsic_computations_mask = scene.comp_mask
        d.intrinsic_free_mask = scene.free_mask
        d.intrinsic_defaults = dptr(arr(scene.intrinsic_defaults, np.float64))
        d.fix_mask = scene.fix_mask
        d.fix_points = dptr(arr(scene.fix_points, np.float64))
        d.fix_images = dptr(arr(scene.fix_images, np.float64))
        d.fix_cameras = dptr(arr(scene.fix_cameras, np.float64))

        def cons(a):
            if a is None or len(a) == 0:
                return 0, C.cast(None, C.POINTER(hsba_constraint))
            buf = (hsba_constraint * len(a))()
            for j, (cid, m) in enumerate(a):
                buf[j].id, buf[j].mask = int(cid), int(m)
            keep.append(buf)
            return len(a), buf

        d.num_point_constraints, d.point_constraints = cons(scene.point_constraints)
        d.num_image_constraints, d.image_constraints = cons(scene.image_constraints)
        d.num_camera_constraints, d.camera_constraints = cons(scene.camera_constraints)
        self.desc = d
        self.kb = arr(scene.key_blocks.reshape(-1) if scene.key_blocks is not None else None,
                      np.float64)
        self.cw = arr(scene.constraint_weights, np.float64)
        pb = getattr(scene, "point_constraint_blocks", None)
        self.pb = arr(None if pb is None else np.asarray(pb).reshape(-1), np.float64)
        self.ycov = hsba_ycov_inverse(dptr(self.kb), dptr(self.cw), dptr(self.pb))
        self._keep = keep


_LIB = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libhsba.so; raises if it is missing (no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"libhsba.so not built at {path}: run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.hsba_version.restype = C.c_char_p
    lib.hsba_device_count.argtypes = [C.POINTER(C.c_int)]
    lib.hsba_problem_create.argtypes = [C.POINTER(hsba_problem_desc), C.POINTER(hsba_options),
                                        C.POINTER(P)]
    lib.hsba_problem_destroy.argtypes = [P]
    lib.hsba_problem_sizes.argtypes = [P] + [C.POINTER(C.c_int64)] * 4
    lib.hsba_evaluate.argtypes = [P, _dp, _dp]
    lib.hsba_jacobian.argtypes = [P, _dp, _dp, _dp, _dp]
    lib.hsba_build_normal.argtypes = [P, _dp, _dp, C.POINTER(hsba_ycov_inverse), _dp, _dp, _dp, _dp]
    lib.hsba_solve.argtypes = [P, C.c_double, _dp]
    lib.hsba_reduced_system.argtypes = [P, C.c_double, _dp, _dp]
    lib.hsba_max_diagonal.argtypes = [P, _dp]
    lib.hsba_gradient_inf_norm.argtypes = [P, _dp]
    lib.hsba_gradient_dot.argtypes = [P, _dp, _dp]
    lib.hsba_mahalanobis.argtypes = [P, _dp, _dp, C.POINTER(hsba_ycov_inverse), _dp]
    lib.hsba_lm_solve.argtypes = [P, _dp, _dp, C.POINTER(hsba_ycov_inverse),
                                  C.POINTER(hsba_lm_options), _dp, C.POINTER(hsba_lm_stats)]
    lib.hsba_lm_begin.argtypes = [P, _dp, _dp, C.POINTER(hsba_ycov_inverse),
                                  C.POINTER(hsba_lm_options)]
    lib.hsba_lm_iterate.argtypes = [P, C.c_int32, C.POINTER(hsba_stage_times)]
    lib.hsba_lm_end.argtypes = [P, _dp, C.POINTER(hsba_lm_stats)]
    lib.hsba_batch_create.argtypes = [C.POINTER(hsba_problem_desc), C.c_int64,
                                      C.POINTER(hsba_options), C.POINTER(P)]
    lib.hsba_batch_destroy.argtypes = [P]
    lib.hsba_batch_lm_solve.argtypes = [P, C.POINTER(_dp), C.POINTER(_dp),
                                        C.POINTER(hsba_ycov_inverse), C.POINTER(hsba_lm_options),
                                        C.POINTER(_dp), C.POINTER(hsba_lm_stats)]
    lib.hsba_server_create.argtypes = [C.POINTER(hsba_server_options), C.POINTER(P)]
    lib.hsba_server_destroy.argtypes = [P]
    lib.hsba_server_submit.argtypes = [P, C.POINTER(hsba_problem_desc), _dp, _dp,
                                       C.POINTER(hsba_ycov_inverse), _dp, C.POINTER(C.c_int64)]
    lib.hsba_server_poll.argtypes = [P, C.POINTER(hsba_job_result), C.c_int32,
                                     C.POINTER(C.c_int32)]
    lib.hsba_server_drain.argtypes = [P]
    lib.hsba_server_clock.argtypes = [P]
    lib.hsba_server_clock.restype = C.c_double
    lib.hsba_naive_problem_create.argtypes = [C.POINTER(hsba_naive_desc), C.POINTER(hsba_options),
                                              C.POINTER(P)]
    lib.hsba_naive_problem_destroy.argtypes = [P]
    lib.hsba_naive_evaluate.argtypes = [P, _dp, _dp]
    lib.hsba_naive_jacobian.argtypes = [P, _dp, _dp, _dp]
    lib.hsba_naive_build_normal.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.hsba_naive_solve.argtypes = [P, C.c_double, _dp]
    lib.hsba_naive_lm_solve.argtypes = [P, _dp, _dp, _dp, C.POINTER(hsba_lm_options), _dp,
                                        C.POINTER(hsba_lm_stats)]
    lib.hsba_gcp_problem_create.argtypes = [C.POINTER(hsba_gcp_desc), C.POINTER(hsba_options),
                                            C.POINTER(P)]
    lib.hsba_gcp_problem_destroy.argtypes = [P]
    lib.hsba_gcp_evaluate.argtypes = [P, _dp, _dp]
    lib.hsba_gcp_jacobian.argtypes = [P, _dp, _dp, _dp]
    lib.hsba_gcp_build_normal.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.hsba_gcp_solve.argtypes = [P, C.c_double, _dp]
    lib.hsba_gcp_mahalanobis.argtypes = [P, _dp, _dp, _dp, _dp, _dp]
    lib.hsba_gcp_lm_solve.argtypes = [P, _dp, _dp, _dp, _dp, C.POINTER(hsba_lm_options), _dp,
                                      C.POINTER(hsba_lm_stats)]
    lib.hsba_tuning_table.argtypes = [C.POINTER(hsba_tune_entry), C.c_int32, C.POINTER(C.c_int32)]
    lib.hsba_tuning_reset.argtypes = []
    lib.hsba_problem_structure.argtypes = [P, C.POINTER(hsba_structure_sizes), _ip, _ip, _ip, _ip, _ip,
                                           _ip, C.POINTER(C.c_int64), _ip]
    lib.hsba_coupling_blocks.argtypes = [P, _dp, _dp, _dp]
    lib.hsba_covariance.argtypes = [P, _dp, _dp, C.POINTER(hsba_ycov_inverse), _dp, _dp, _dp, _dp]
    lib.hsba_covariance_static.argtypes = [P, _dp, _dp, C.POINTER(hsba_ycov_inverse), _dp, _dp, _dp, _dp]
    lib.hsba_x_covariance.argtypes = [P, _dp, _dp, C.POINTER(hsba_ycov_inverse), C.c_int32, _dp,
                                      C.POINTER(hsba_covariance_info)]
    lib.hsba_batch_lm_begin.argtypes = [P, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(hsba_ycov_inverse),
                                        C.POINTER(hsba_lm_options)]
    lib.hsba_batch_lm_iterate.argtypes = [P, C.c_int32, C.POINTER(hsba_stage_times)]
    lib.hsba_batch_lm_end.argtypes = [P, C.POINTER(_dp), C.POINTER(hsba_lm_stats)]
    lib.hsba_scene_load_sfm.argtypes = [C.POINTER(hsba_sfm_files), C.c_uint32, C.c_uint32, C.POINTER(P)]
    lib.hsba_scene_load_bal.argtypes = [C.c_char_p, C.POINTER(P)]
    lib.hsba_scene_get.argtypes = [P, C.POINTER(hsba_scene_view)]
    lib.hsba_scene_store_sfm.argtypes = [P, _dp, C.c_char_p, C.c_char_p, C.c_char_p]
    lib.hsba_scene_store_bal.argtypes = [P, _dp, C.c_char_p]
    lib.hsba_scene_destroy.argtypes = [P]
    _LIB = lib
    return lib


EXPORTED_SYMBOLS = [
    "hsba_version", "hsba_device_count", "hsba_problem_create", "hsba_problem_destroy",
    "hsba_problem_sizes", "hsba_evaluate", "hsba_jacobian", "hsba_build_normal",
    "hsba_solve", "hsba_covariance", "hsba_covariance_static", "hsba_x_covariance", "hsba_problem_structure", "hsba_coupling_blocks",
    "hsba_tuning_table", "hsba_tuning_reset", "hsba_reduced_system", "hsba_max_diagonal", "hsba_gradient_inf_norm",
    "hsba_gradient_dot", "hsba_mahalanobis", "hsba_lm_solve", "hsba_lm_begin",
    "hsba_lm_iterate", "hsba_lm_end", "hsba_batch_create",
    "hsba_batch_destroy", "hsba_batch_lm_solve", "hsba_batch_lm_begin", "hsba_batch_lm_iterate",
    "hsba_batch_lm_end", "hsba_server_create",
    "hsba_server_destroy", "hsba_server_submit", "hsba_server_poll", "hsba_server_drain",
    "hsba_server_clock", "hsba_naive_problem_create", "hsba_naive_problem_destroy",
    "hsba_naive_evaluate", "hsba_naive_jacobian", "hsba_naive_build_normal", "hsba_naive_solve",
    "hsba_naive_lm_solve", "hsba_gcp_problem_create", "hsba_gcp_problem_destroy", "hsba_gcp_evaluate",
    "hsba_gcp_jacobian", "hsba_gcp_build_normal", "hsba_gcp_solve", "hsba_gcp_mahalanobis",
    "hsba_gcp_lm_solve", "hsba_scene_load_sfm", "hsba_scene_load_bal", "hsba_scene_get",
    "hsba_scene_store_sfm", "hsba_scene_store_bal", "hsba_scene_destroy",
]


def check(rc: int, what: str = ""):
    if rc != HSBA_OK:
        raise RuntimeError(f"{what} failed: {rc} ({ERRORS.get(rc, '?')})")


def tuning_table():
    """The autotuner's decisions: list of dicts (precision, camera block E,
    incidence bucket, chosen pair kernel, measured ms of each candidate)."""
    lib = load_library()
    n = C.c_int32()
    check(lib.hsba_tuning_table(None, 0, C.byref(n)), "hsba_tuning_table")
    arr = (hsba_tune_entry * max(1, n.value))()
    check(lib.hsba_tuning_table(arr, n.value, C.byref(n)), "hsba_tuning_table")
    names = ["valu_wave_per_pair", "mfma_workgroup_per_pair", "mfma_wave_per_pair"]
    return [{"precision": "fp64" if e.precision == PRECISION_FP64 else "mixed", "camera_block": e.camera_block,
             "incidences_per_pair_bucket": f"[{2 ** e.bucket}, {2 ** (e.bucket + 1)})",
             "chosen": names[e.variant], "ms": {names[v]: e.ms[v] for v in range(3)}, "pairs_timed": e.pairs}
            for e in arr[:n.value]]
