"""Python host mirror of the reference's bundle-adjustment interface over the
C ABI of libhsba.so (include/hsba.h).

Class names follow hs_sfm's camera_shared family so tests read like the
reference's own:

  CameraSharedVectorFunction             operator()(x) -> y
      (include/hs_sfm/bundle_adjustment/camera_shared_vector_function.hpp:412-632)
  CameraSharedAnalyticalJacobianMatrixCalculator   (…_analytical_jacobian_matrix_calculator.hpp:66-332)
  CameraSharedNormalEquationBuilder      (…_normal_equation_builder.hpp:92-462)
  CameraSharedNormalEquationSolver       (…_normal_equation_solver.hpp:75-114)
  CameraSharedLevenbergMarquardtOptimizor(…_levenberg_marquardt_optimizor.hpp:19-63)

Errors: the reference returns `Err` 0 / -1; these wrappers return the
library's code (0 or negative) and raise only on misuse.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import abi
from .synthetic import Scene

_dp = C.POINTER(C.c_double)


def _opts(precision=abi.PRECISION_MIXED, solver=abi.SOLVER_AUTO, pcg_rtol=1e-6,
          pcg_max_iterations=500, point_group_lanes=0, loss=abi.LOSS_TRIVIAL, loss_scale=1.0,
          pair_kernel=abi.PAIRS_TUNED):
    return abi.hsba_options(precision=precision, solver=solver, pcg_rtol=pcg_rtol,
                            pcg_max_iterations=pcg_max_iterations,
                            point_group_lanes=point_group_lanes, loss=loss, loss_scale=loss_scale,
                            pair_kernel=pair_kernel)


class Problem:
    """Device-resident problem (hsba_problem) built from a Scene."""

    def __init__(self, scene: Scene, precision=abi.PRECISION_MIXED, solver=abi.SOLVER_AUTO,
                 pcg_rtol=1e-6, point_group_lanes=0, loss=abi.LOSS_TRIVIAL, loss_scale=1.0,
                 pair_kernel=abi.PAIRS_TUNED):
        self.lib = abi.load_library()
        self.scene = scene
        self.desc = abi.Desc(scene)
        self.h = C.c_void_p()
        o = _opts(precision, solver, pcg_rtol, point_group_lanes=point_group_lanes, loss=loss,
                  loss_scale=loss_scale, pair_kernel=pair_kernel)
        rc = self.lib.hsba_problem_create(C.byref(self.desc.desc), C.byref(o), C.byref(self.h))
        abi.check(rc, "hsba_problem_create")
        xs, ys, dd, nr = (C.c_int64() for _ in range(4))
        self.lib.hsba_problem_sizes(self.h, C.byref(xs), C.byref(ys), C.byref(dd), C.byref(nr))
        self.x_size, self.y_size, self.cam_d, self.n_reduced = xs.value, ys.value, dd.value, nr.value

    def close(self):
        if self.h:
            self.lib.hsba_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _d(a):
        return abi.dptr(np.ascontiguousarray(a, np.float64))

    def evaluate(self, x):
        y = np.zeros(self.y_size)
        abi.check(self.lib.hsba_evaluate(self.h, self._d(x), abi.dptr(y)), "hsba_evaluate")
        return y

    def jacobian(self, x):
        N = self.scene.num_keys
        Jp, Ji = np.zeros((N, 2, 3)), np.zeros((N, 2, 6))
        Jc = np.zeros((N, 2, max(1, self.cam_d)))
        abi.check(self.lib.hsba_jacobian(self.h, self._d(x), abi.dptr(Jp), abi.dptr(Ji),
                                         abi.dptr(Jc)), "hsba_jacobian")
        return Jp, Ji, Jc[:, :, :self.cam_d]

    def build_normal(self, x, y_obs=None):
        s = self.scene
        y_obs = s.y_obs if y_obs is None else y_obs
        d = self.cam_d
        V = np.zeros((s.num_points, 3, 3))
        Ui = np.zeros((s.num_images, 6, 6))
        Uc = np.zeros((s.num_cameras, max(d, 1), max(d, 1)))
        g = np.zeros(self.x_size)
        rc = self.lib.hsba_build_normal(self.h, self._d(x), self._d(y_obs), C.byref(self.desc.ycov),
                                        abi.dptr(V), abi.dptr(Ui), abi.dptr(Uc), abi.dptr(g))
        abi.check(rc, "hsba_build_normal")
        return dict(V=V, Ui=Ui, Uc=Uc[:, :d, :d] if d else Uc[:, :0, :0], g=g)

    def solve(self, mu):
        delta = np.zeros(self.x_size)
        rc = self.lib.hsba_solve(self.h, mu, abi.dptr(delta))
        return rc, delta

    def covariance(self, x=None, y_obs=None, static=False):
        """Diagonal blocks of N^-1 at x and the dense inverse of the reduced
        system (hsba_covariance, or hsba_covariance_static with static=True:
        no device allocation); returns (rc, dict)."""
        s = self.scene
        x = s.x0 if x is None else x
        y_obs = s.y_obs if y_obs is None else y_obs
        d = self.cam_d
        out = dict(V=np.zeros((s.num_points, 3, 3)), Ui=np.zeros((s.num_images, 6, 6)),
                   Uc=np.zeros((s.num_cameras, max(d, 1), max(d, 1))),
                   reduced=np.zeros((self.n_reduced, self.n_reduced)))
        fn = self.lib.hsba_covariance_static if static else self.lib.hsba_covariance
        rc = fn(self.h, self._d(x), self._d(y_obs), C.byref(self.desc.ycov), abi.dptr(out["V"]),
                abi.dptr(out["Ui"]), abi.dptr(out["Uc"]), abi.dptr(out["reduced"]))
        out["Uc"] = out["Uc"][:, :d, :d]
        return rc, out

    def x_covariance(self, x=None, y_obs=None, method=abi.COV_AUTO):
        """hsba_x_covariance: the full dense covariance [x_size][x_size] at x
        (the pseudo-inverse of N with singular values <= 1e-6 dropped);
        returns (rc, C, info dict)."""
        s = self.scene
        x = s.x0 if x is None else x
        y_obs = s.y_obs if y_obs is None else y_obs
        out = np.zeros((self.x_size, self.x_size))
        info = abi.hsba_covariance_info()
        rc = self.lib.hsba_x_covariance(self.h, self._d(x), self._d(y_obs), C.byref(self.desc.ycov), method,
                                        abi.dptr(out), C.byref(info))
        return rc, out, dict(method=info.method, rank=info.rank, sweeps=info.sweeps,
                             smallest_kept=info.smallest_kept)

    def reduced_system(self, mu):
        n = self.n_reduced
        S, rhs = np.zeros((n, n)), np.zeros(n)
        rc = self.lib.hsba_reduced_system(self.h, mu, abi.dptr(S), abi.dptr(rhs))
        return rc, S, rhs

    def structure(self):
        """hsba_problem_structure: the index structure as numpy arrays."""
        sz = abi.hsba_structure_sizes()
        abi.check(self.lib.hsba_problem_structure(self.h, C.byref(sz), *([abi.iptr(None)] * 6),
                                                  C.cast(None, C.POINTER(C.c_int64)), abi.iptr(None)),
                  "hsba_problem_structure")
        out = dict(key_order=np.zeros(sz.num_keys, np.int32), point_ptr=np.zeros(sz.num_points + 1, np.int32),
                   point_image=np.zeros((sz.num_point_image, 2), np.int32),
                   point_camera=np.zeros((sz.num_point_camera, 2), np.int32),
                   image_camera=np.zeros((sz.num_image_camera, 2), np.int32),
                   pairs=np.zeros((sz.num_pairs, 2), np.int32), pair_ptr=np.zeros(sz.num_pairs + 1, np.int64),
                   incidences=np.zeros((sz.num_incidences, 2), np.int32))
        abi.check(self.lib.hsba_problem_structure(
            self.h, C.byref(sz), abi.iptr(out["key_order"]), abi.iptr(out["point_ptr"]),
            abi.iptr(out["point_image"]), abi.iptr(out["point_camera"]), abi.iptr(out["image_camera"]),
            abi.iptr(out["pairs"]), out["pair_ptr"].ctypes.data_as(C.POINTER(C.c_int64)),
            abi.iptr(out["incidences"])), "hsba_problem_structure")
        return out

    def coupling_blocks(self):
        """hsba_coupling_blocks of the last build: W_pi, W_pc, W_ic."""
        st = self.structure()
        d = self.cam_d
        Wpi = np.zeros((len(st["point_image"]), 3, 6))
        Wpc = np.zeros((len(st["point_camera"]), 3, max(d, 1)))
        Wic = np.zeros((len(st["image_camera"]), 6, max(d, 1)))
        abi.check(self.lib.hsba_coupling_blocks(self.h, abi.dptr(Wpi), abi.dptr(Wpc), abi.dptr(Wic)),
                  "hsba_coupling_blocks")
        return Wpi, Wpc[:, :, :d], Wic[:, :, :d]

    def max_diagonal(self):
        v = C.c_double()
        abi.check(self.lib.hsba_max_diagonal(self.h, C.byref(v)), "hsba_max_diagonal")
        return v.value

    def gradient_inf_norm(self):
        v = C.c_double()
        abi.check(self.lib.hsba_gradient_inf_norm(self.h, C.byref(v)), "hsba_gradient_inf_norm")
        return v.value

    def gradient_dot(self, delta):
        v = C.c_double()
        abi.check(self.lib.hsba_gradient_dot(self.h, self._d(delta), C.byref(v)), "hsba_gradient_dot")
        return v.value

    def mahalanobis(self, x, y_obs=None):
        y_obs = self.scene.y_obs if y_obs is None else y_obs
        v = C.c_double()
        abi.check(self.lib.hsba_mahalanobis(self.h, self._d(x), self._d(y_obs),
                                            C.byref(self.desc.ycov), C.byref(v)), "hsba_mahalanobis")
        return v.value

    def lm(self, x0=None, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0):
        s = self.scene
        x0 = s.x0 if x0 is None else x0
        opt = abi.hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        x = np.zeros(self.x_size)
        st = abi.hsba_lm_stats()
        rc = self.lib.hsba_lm_solve(self.h, self._d(x0), self._d(s.y_obs), C.byref(self.desc.ycov),
                                    C.byref(opt), abi.dptr(x), C.byref(st))
        return rc, x, st


    # stepwise LM (timing / serving)
    def lm_begin(self, x0=None, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0):
        s = self.scene
        self._x0 = np.ascontiguousarray(s.x0 if x0 is None else x0, np.float64)
        self._y = np.ascontiguousarray(s.y_obs, np.float64)
        opt = abi.hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        return self.lib.hsba_lm_begin(self.h, abi.dptr(self._x0), abi.dptr(self._y),
                                      C.byref(self.desc.ycov), C.byref(opt))

    def lm_iterate(self, iterations, timed=False):
        t = abi.hsba_stage_times() if timed else None
        rc = self.lib.hsba_lm_iterate(self.h, iterations, C.byref(t) if timed else None)
        return rc, (t.as_dict() if timed else None)

    def lm_end(self):
        x = np.zeros(self.x_size)
        st = abi.hsba_lm_stats()
        rc = self.lib.hsba_lm_end(self.h, abi.dptr(x), C.byref(st))
        return rc, x, st


class Batch:
    """Many independent windows solved by one set of launches per LM iteration."""

    def __init__(self, scenes: Sequence[Scene], precision=abi.PRECISION_MIXED, point_group_lanes=0,
                 pair_kernel=abi.PAIRS_TUNED):
        self.lib = abi.load_library()
        self.scenes = list(scenes)
        self.descs = [abi.Desc(s) for s in self.scenes]
        n = len(self.scenes)
        self.desc_arr = (abi.hsba_problem_desc * n)(*[d.desc for d in self.descs])
        self.ycov_arr = (abi.hsba_ycov_inverse * n)(*[d.ycov for d in self.descs])
        self.h = C.c_void_p()
        o = _opts(precision, point_group_lanes=point_group_lanes, pair_kernel=pair_kernel)
        rc = self.lib.hsba_batch_create(self.desc_arr, n, C.byref(o), C.byref(self.h))
        abi.check(rc, "hsba_batch_create")

    def close(self):
        if self.h:
            self.lib.hsba_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lm(self, x0s=None, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0):
        n = len(self.scenes)
        x0s = [s.x0 for s in self.scenes] if x0s is None else x0s
        self._x0 = [np.ascontiguousarray(x, np.float64) for x in x0s]
        self._y = [np.ascontiguousarray(s.y_obs, np.float64) for s in self.scenes]
        xs = [np.zeros(s.x_size) for s in self.scenes]
        X0 = (_dp * n)(*[abi.dptr(a) for a in self._x0])
        Y = (_dp * n)(*[abi.dptr(a) for a in self._y])
        XO = (_dp * n)(*[abi.dptr(a) for a in xs])
        st = (abi.hsba_lm_stats * n)()
        opt = abi.hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        rc = self.lib.hsba_batch_lm_solve(self.h, X0, Y, self.ycov_arr, C.byref(opt), XO, st)
        return rc, xs, list(st)

    def lm_begin(self, x0s=None, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0):
        """Stepwise batch solve (hsba_batch_lm_begin / _iterate / _end)."""
        n = len(self.scenes)
        x0s = [s.x0 for s in self.scenes] if x0s is None else x0s
        self._x0 = [np.ascontiguousarray(x, np.float64) for x in x0s]
        self._y = [np.ascontiguousarray(s.y_obs, np.float64) for s in self.scenes]
        X0 = (_dp * n)(*[abi.dptr(a) for a in self._x0])
        Y = (_dp * n)(*[abi.dptr(a) for a in self._y])
        opt = abi.hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        return self.lib.hsba_batch_lm_begin(self.h, X0, Y, self.ycov_arr, C.byref(opt))

    def lm_iterate(self, iterations, timed=False):
        t = abi.hsba_stage_times() if timed else None
        rc = self.lib.hsba_batch_lm_iterate(self.h, iterations, C.byref(t) if timed else None)
        return rc, (t.as_dict() if timed else None)

    def lm_end(self):
        n = len(self.scenes)
        xs = [np.zeros(s.x_size) for s in self.scenes]
        XO = (_dp * n)(*[abi.dptr(a) for a in xs])
        st = (abi.hsba_lm_stats * n)()
        rc = self.lib.hsba_batch_lm_end(self.h, XO, st)
        return rc, xs, list(st)


class NaiveProblem:
    """ba_naive problem over the C ABI: x = [cameras 6C | points 3P], features
    in normalised coordinates (ba_naive_vector_function.hpp:117-196)."""

    def __init__(self, num_cameras, num_points, feature_maps, precision=abi.PRECISION_FP64):
        self.lib = abi.load_library()
        self.C, self.P = int(num_cameras), int(num_points)
        self.fm = np.ascontiguousarray(feature_maps, np.int32)
        self.F = self.fm.shape[0]
        d = abi.hsba_naive_desc(self.C, self.P, self.F, self.fm.ctypes.data_as(C.POINTER(C.c_int32)))
        self.h = C.c_void_p()
        o = _opts(precision)
        self.last_rc = self.lib.hsba_naive_problem_create(C.byref(d), C.byref(o), C.byref(self.h))
        abi.check(self.last_rc, "hsba_naive_problem_create")
        self.x_size = 6 * self.C + 3 * self.P

    def close(self):
        if self.h:
            self.lib.hsba_naive_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(2 * self.F)
        abi.check(self.lib.hsba_naive_evaluate(self.h, abi.dptr(x), abi.dptr(y)), "hsba_naive_evaluate")
        return y

    def jacobian(self, x):
        """Returns (rc, camera blocks [F,2,6], point blocks [F,2,3])."""
        x = np.ascontiguousarray(x, np.float64)
        jc, jp = np.zeros((self.F, 2, 6)), np.zeros((self.F, 2, 3))
        rc = self.lib.hsba_naive_jacobian(self.h, abi.dptr(x), abi.dptr(jc), abi.dptr(jp))
        return rc, jc, jp

    def build_normal(self, x, y_obs, cov=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y_obs, np.float64)
        cov = None if cov is None else np.ascontiguousarray(cov, np.float64)
        U, V, g = np.zeros((self.C, 6, 6)), np.zeros((self.P, 3, 3)), np.zeros(self.x_size)
        rc = self.lib.hsba_naive_build_normal(self.h, abi.dptr(x), abi.dptr(y), abi.dptr(cov), abi.dptr(U),
                                              abi.dptr(V), abi.dptr(g))
        abi.check(rc, "hsba_naive_build_normal")
        return U, V, g

    def solve(self, mu):
        dx = np.zeros(self.x_size)
        rc = self.lib.hsba_naive_solve(self.h, mu, abi.dptr(dx))
        return rc, dx

    def lm(self, x0, y_obs, cov=None, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0):
        x0 = np.ascontiguousarray(x0, np.float64)
        y = np.ascontiguousarray(y_obs, np.float64)
        cov = None if cov is None else np.ascontiguousarray(cov, np.float64)
        xo = np.zeros(self.x_size)
        st = abi.hsba_lm_stats()
        opt = abi.hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        rc = self.lib.hsba_naive_lm_solve(self.h, abi.dptr(x0), abi.dptr(y), abi.dptr(cov), C.byref(opt),
                                          abi.dptr(xo), C.byref(st))
        return rc, xo, st


class GCPProblem:
    """ba_gcp_constrained problem over the C ABI: ba_naive whose last
    num_gcps points carry observed coordinates with 3x3 inverse covariances;
    y = [features 2F | GCP coordinates 3G]
    (ba_gcp_constrained_vector_function.hpp:73-113)."""

    def __init__(self, num_cameras, num_points, feature_maps, num_gcps, precision=abi.PRECISION_FP64):
        self.lib = abi.load_library()
        self.C, self.P, self.G = int(num_cameras), int(num_points), int(num_gcps)
        self.fm = np.ascontiguousarray(feature_maps, np.int32)
        self.F = self.fm.shape[0]
        d = abi.hsba_gcp_desc(self.C, self.P, self.F, self.fm.ctypes.data_as(C.POINTER(C.c_int32)), self.G)
        self.h = C.c_void_p()
        o = _opts(precision)
        self.last_rc = self.lib.hsba_gcp_problem_create(C.byref(d), C.byref(o), C.byref(self.h))
        abi.check(self.last_rc, "hsba_gcp_problem_create")
        self.x_size = 6 * self.C + 3 * self.P
        self.y_size = 2 * self.F + 3 * self.G

    def close(self):
        if self.h:
            self.lib.hsba_gcp_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _arr(a):
        return None if a is None else np.ascontiguousarray(a, np.float64)

    def evaluate(self, x):
        y = np.zeros(self.y_size)
        abi.check(self.lib.hsba_gcp_evaluate(self.h, abi.dptr(self._arr(x)), abi.dptr(y)), "hsba_gcp_evaluate")
        return y

    def jacobian(self, x):
        """Returns (rc, camera blocks [F,2,6], point blocks [F,2,3])."""
        jc, jp = np.zeros((self.F, 2, 6)), np.zeros((self.F, 2, 3))
        rc = self.lib.hsba_gcp_jacobian(self.h, abi.dptr(self._arr(x)), abi.dptr(jc), abi.dptr(jp))
        return rc, jc, jp

    def build_normal(self, x, y_obs, cov=None, gcp_blocks=None):
        U, V, g = np.zeros((self.C, 6, 6)), np.zeros((self.P, 3, 3)), np.zeros(self.x_size)
        rc = self.lib.hsba_gcp_build_normal(self.h, abi.dptr(self._arr(x)), abi.dptr(self._arr(y_obs)),
                                            abi.dptr(self._arr(cov)), abi.dptr(self._arr(gcp_blocks)),
                                            abi.dptr(U), abi.dptr(V), abi.dptr(g))
        abi.check(rc, "hsba_gcp_build_normal")
        return U, V, g

    def solve(self, mu):
        dx = np.zeros(self.x_size)
        rc = self.lib.hsba_gcp_solve(self.h, mu, abi.dptr(dx))
        return rc, dx

    def mahalanobis(self, x, y_obs, cov=None, gcp_blocks=None):
        v = C.c_double()
        rc = self.lib.hsba_gcp_mahalanobis(self.h, abi.dptr(self._arr(x)), abi.dptr(self._arr(y_obs)),
                                           abi.dptr(self._arr(cov)), abi.dptr(self._arr(gcp_blocks)), C.byref(v))
        abi.check(rc, "hsba_gcp_mahalanobis")
        return v.value

    def lm(self, x0, y_obs, cov=None, gcp_blocks=None, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0):
        xo = np.zeros(self.x_size)
        st = abi.hsba_lm_stats()
        opt = abi.hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        rc = self.lib.hsba_gcp_lm_solve(self.h, abi.dptr(self._arr(x0)), abi.dptr(self._arr(y_obs)),
                                        abi.dptr(self._arr(cov)), abi.dptr(self._arr(gcp_blocks)), C.byref(opt),
                                        abi.dptr(xo), C.byref(st))
        return rc, xo, st


class SceneFile:
    """A problem read from files through the C ABI (hsba_scene_*): the hs_sfm
    text formats (sfm_file_io/*_loader.hpp, assembled as
    BundleAdjustmentOptimizor does) or a BAL file.  `scene` is a Scene ready
    for Problem / Batch / Server; store_*() write an optimised x back."""

    def __init__(self, handle, lib):
        self.lib, self.h = lib, handle
        v = abi.hsba_scene_view()
        abi.check(lib.hsba_scene_get(self.h, C.byref(v)), "hsba_scene_get")
        d = v.desc
        N, I, Cc = d.num_keys, d.num_images, d.num_cameras
        fm = np.ctypeslib.as_array(d.feature_maps, (max(N, 1), 2))[:N].copy() if N else np.zeros((0, 2), np.int32)
        icm = np.ctypeslib.as_array(d.image_camera_map, (max(I, 1),))[:I].copy()
        intr = np.ctypeslib.as_array(d.intrinsic_defaults, (max(Cc, 1), 10))[:Cc].copy()
        self.scene = Scene(num_points=int(d.num_points), num_images=int(I), num_cameras=int(Cc),
                           feature_maps=fm.astype(np.int32), image_camera_map=icm.astype(np.int32),
                           comp_mask=int(d.intrinsic_computations_mask), free_mask=int(d.intrinsic_free_mask),
                           intrinsic_defaults=intr, fix_mask=0, name="file")
        self.scene.x0 = np.ctypeslib.as_array(v.x0, (max(v.x_size, 1),))[:v.x_size].copy()
        self.scene.y_obs = np.ctypeslib.as_array(v.y_obs, (max(v.y_size, 1),))[:v.y_size].copy()

    @classmethod
    def load_sfm(cls, tracks, keysets, intrinsics, extrinsics, points, image_intrinsic_map,
                 image_extrinsic_map, track_point_map, comp_mask=7, free_mask=0):
        lib = abi.load_library()
        enc = [None if k is None else str(k).encode() for k in keysets]
        ks = (C.c_char_p * max(1, len(enc)))(*enc)
        f = abi.hsba_sfm_files(str(tracks).encode(), ks, len(enc), str(intrinsics).encode(),
                               str(extrinsics).encode(), str(points).encode(),
                               str(image_intrinsic_map).encode(), str(image_extrinsic_map).encode(),
                               str(track_point_map).encode())
        h = C.c_void_p()
        abi.check(lib.hsba_scene_load_sfm(C.byref(f), comp_mask, free_mask, C.byref(h)), "hsba_scene_load_sfm")
        return cls(h, lib)

    @classmethod
    def load_bal(cls, path):
        lib = abi.load_library()
        h = C.c_void_p()
        abi.check(lib.hsba_scene_load_bal(str(path).encode(), C.byref(h)), "hsba_scene_load_bal")
        return cls(h, lib)

    def store_sfm(self, x, intrinsics=None, extrinsics=None, points=None):
        enc = [None if p is None else str(p).encode() for p in (intrinsics, extrinsics, points)]
        abi.check(self.lib.hsba_scene_store_sfm(self.h, abi.dptr(np.ascontiguousarray(x, np.float64)), *enc),
                  "hsba_scene_store_sfm")

    def store_bal(self, x, path):
        abi.check(self.lib.hsba_scene_store_bal(self.h, abi.dptr(np.ascontiguousarray(x, np.float64)),
                                                str(path).encode()), "hsba_scene_store_bal")

    def close(self):
        if self.h:
            self.lib.hsba_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over the C ABI (continuous batching of independent jobs).

    submit() copies a job in and returns its id at once; poll() returns the
    jobs finished since the last poll; drain() waits for every submitted job.
    Each job's result is written into the array returned by submit()."""

    def __init__(self, precision=abi.PRECISION_MIXED, hbm_budget_bytes=0, max_active_jobs=0,
                 streams=0, max_batch_jobs=0, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0,
                 loss=abi.LOSS_TRIVIAL, loss_scale=1.0):
        self.lib = abi.load_library()
        o = abi.hsba_server_options(hbm_budget_bytes=hbm_budget_bytes, max_active_jobs=max_active_jobs,
                                    precision=precision, streams=streams, max_batch_jobs=max_batch_jobs,
                                    loss=loss, loss_scale=loss_scale)
        o.lm = abi.hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        self.h = C.c_void_p()
        abi.check(self.lib.hsba_server_create(C.byref(o), C.byref(self.h)), "hsba_server_create")
        self._jobs = {}  # job id -> (desc holder, x0, y, x_out)

    def prepare(self, scene: Scene, x0=None):
        """Builds the job's C structures ahead of submission (no library call)."""
        d = abi.Desc(scene)
        x0 = np.ascontiguousarray(scene.x0 if x0 is None else x0, np.float64)
        y = np.ascontiguousarray(scene.y_obs, np.float64)
        return (d, x0, y, np.zeros(scene.x_size))

    def submit(self, job) -> int:
        d, x0, y, xo = job
        jid = C.c_int64()
        rc = self.lib.hsba_server_submit(self.h, C.byref(d.desc), abi.dptr(x0), abi.dptr(y),
                                         C.byref(d.ycov), abi.dptr(xo), C.byref(jid))
        abi.check(rc, "hsba_server_submit")
        self._jobs[jid.value] = job
        return jid.value

    def x_out(self, job_id: int) -> np.ndarray:
        return self._jobs[job_id][3]

    def poll(self, max_results=1024):
        res = (abi.hsba_job_result * max_results)()
        n = C.c_int32()
        abi.check(self.lib.hsba_server_poll(self.h, res, max_results, C.byref(n)), "hsba_server_poll")
        return [res[i] for i in range(n.value)]

    def drain(self):
        abi.check(self.lib.hsba_server_drain(self.h), "hsba_server_drain")

    def clock(self) -> float:
        return self.lib.hsba_server_clock(self.h)

    def close(self):
        if self.h:
            self.lib.hsba_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Reference-named wrappers ----------------------------------------------------
class CameraSharedVectorFunction:
    def __init__(self, problem: Problem):
        self.problem = problem

    def __call__(self, x):
        return 0, self.problem.evaluate(x)


class CameraSharedAnalyticalJacobianMatrixCalculator:
    def __call__(self, vector_function: CameraSharedVectorFunction, x):
        return 0, vector_function.problem.jacobian(x)


class CameraSharedNormalEquationBuilder:
    def __call__(self, problem: Problem, x, y_obs):
        return 0, problem.build_normal(x, y_obs)


class CameraSharedNormalEquationSolver:
    def __call__(self, problem: Problem, mu):
        return problem.solve(mu)


class CameraSharedLevenbergMarquardtOptimizor:
    """ctor(x0, max_itr, tau, eps1, eps2); operator()(problem) -> (err, x)."""

    def __init__(self, initial_x, max_itr_num=20, tau=1e-3, eps1=0.0, eps2=0.0):
        self.x0 = np.asarray(initial_x, np.float64)
        self.max_itr, self.tau, self.eps1, self.eps2 = max_itr_num, tau, eps1, eps2

    def __call__(self, problem: Problem):
        if self.x0.shape[0] != problem.x_size:
            return -1, None
        rc, x, st = problem.lm(self.x0, self.max_itr, self.tau, self.eps1, self.eps2)
        self.stats = st
        return rc, x
