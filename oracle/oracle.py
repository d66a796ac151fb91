"""Python binding of the FP64 CPU oracle (oracle/hsba_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's CPU-baseline leg; never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from hs_sfm_amd.abi import Desc, dptr, hsba_lm_options, hsba_lm_stats, hsba_problem_desc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_dp = C.POINTER(C.c_double)


class LMTrace(C.Structure):
    _fields_ = [("mu", C.c_double), ("cost", C.c_double), ("cost_new", C.c_double),
                ("rho", C.c_double), ("step_norm", C.c_double), ("accepted", C.c_int32),
                ("reserved_", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(hsba_problem_desc), C.POINTER(C.c_int)]
        L.oracle_create.restype = P
        L.oracle_destroy.argtypes = [P]
        L.oracle_sizes.argtypes = [P] + [C.POINTER(C.c_int64)] * 4
        L.oracle_evaluate.argtypes = [P, _dp, _dp]
        L.oracle_jacobian.argtypes = [P, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.oracle_build_normal.argtypes = [P, _dp, _dp, _dp, _dp] + [_dp] * 7
        L.oracle_reduced_system.argtypes = [P, C.c_double, C.c_int, _dp, _dp]
        L.oracle_solve.argtypes = [P, C.c_double, C.c_int, _dp]
        for f in ("oracle_max_diagonal", "oracle_gradient_inf_norm"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = C.c_double
        L.oracle_gradient_dot.argtypes = [P, _dp]
        L.oracle_gradient_dot.restype = C.c_double
        L.oracle_mahalanobis.argtypes = [P, _dp, _dp, _dp, _dp]
        L.oracle_mahalanobis.restype = C.c_double
        L.oracle_lm.argtypes = [P, _dp, _dp, _dp, _dp, C.POINTER(hsba_lm_options), C.c_int, _dp,
                                C.POINTER(hsba_lm_stats), C.POINTER(LMTrace)]
        L.oracle_coupling.argtypes = [P, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int32), _dp]
        L.oracle_pairs.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
        L.oracle_dense_normal.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_set_point_blocks.argtypes = [P, _dp]
        L.oracle_set_loss.argtypes = [P, C.c_int, C.c_double]
        _lib = L
    return _lib


class Oracle:
    """FP64 CPU restatement of the camera_shared LM path for one scene."""

    def __init__(self, scene, loss=0, loss_scale=1.0):
        """loss: 0 none, 1 Huber with scale loss_scale on each key's r^T W r."""
        self.scene = scene
        self.d = Desc(scene)
        rc = C.c_int(0)
        self.h = lib().oracle_create(C.byref(self.d.desc), C.byref(rc))
        if not self.h:
            raise ValueError(f"oracle_create failed: {rc.value}")
        xs, ys, dd, nr = (C.c_int64() for _ in range(4))
        lib().oracle_sizes(self.h, C.byref(xs), C.byref(ys), C.byref(dd), C.byref(nr))
        self.x_size, self.y_size, self.cam_d, self.n_reduced = xs.value, ys.value, dd.value, nr.value
        if self.d.pb is not None:
            rc = lib().oracle_set_point_blocks(self.h, dptr(self.d.pb))
            if rc:
                raise ValueError(f"oracle_set_point_blocks failed: {rc}")
        if loss:
            rc = lib().oracle_set_loss(self.h, int(loss), float(loss_scale))
            if rc:
                raise ValueError(f"oracle_set_loss failed: {rc}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def evaluate(self, x):
        y = np.zeros(self.y_size)
        lib().oracle_evaluate(self.h, dptr(np.ascontiguousarray(x, np.float64)), dptr(y))
        return y

    def jacobian(self, x, mode=0, delta=1e-6, min_delta=1e-8):
        """mode 0 analytic, 1 forward FD, 2 central FD, 3 complex step."""
        N = self.scene.num_keys
        Jp, Ji, Jc = np.zeros((N, 2, 3)), np.zeros((N, 2, 6)), np.zeros((N, 2, max(1, self.cam_d)))
        lib().oracle_jacobian(self.h, dptr(np.ascontiguousarray(x, np.float64)), mode, delta,
                              min_delta, dptr(Jp), dptr(Ji), dptr(Jc))
        return Jp, Ji, Jc[:, :, :self.cam_d]

    def build_normal(self, x, y_obs=None):
        s = self.scene
        y_obs = s.y_obs if y_obs is None else y_obs
        d = self.cam_d
        P, I, Cc, N = s.num_points, s.num_images, s.num_cameras, s.num_keys
        out = dict(V=np.zeros((P, 3, 3)), Ui=np.zeros((I, 6, 6)), Uc=np.zeros((Cc, d, d)),
                   Wpi=np.zeros((N, 3, 6)), Wpc=np.zeros((N, 3, d)), Wic=np.zeros((I, 6, d)),
                   g=np.zeros(self.x_size))
        rc = lib().oracle_build_normal(
            self.h, dptr(np.ascontiguousarray(x, np.float64)),
            dptr(np.ascontiguousarray(y_obs, np.float64)), dptr(self.d.kb), dptr(self.d.cw),
            *(dptr(out[k]) for k in ("V", "Ui", "Uc", "Wpi", "Wpc", "Wic", "g")))
        assert rc == 0
        return out

    def coupling(self, which, blocks=True):
        """Coupling set `which` (0 point-image, 1 point-camera, 2 image-camera)
        in the reference's block-id order, and its blocks at the last build."""
        n = C.c_int64()
        lib().oracle_coupling(self.h, which, C.byref(n), None, None)
        rc = np.zeros((n.value, 2), np.int32)
        rows, cols = [(3, 6), (3, self.cam_d), (6, self.cam_d)][which]
        W = np.zeros((n.value, rows, max(cols, 1))) if blocks else None
        r = lib().oracle_coupling(self.h, which, C.byref(n), rc.ctypes.data_as(C.POINTER(C.c_int32)),
                                  dptr(W) if blocks else None)
        assert r == 0
        return rc, (W[:, :, :cols] if blocks else None)

    def pairs(self):
        """Image pairs sharing a point, their incidence offsets and incidences."""
        n, m = C.c_int64(), C.c_int64()
        lib().oracle_pairs(self.h, C.byref(n), C.byref(m), None, None, None)
        pr, ptr, inc = np.zeros((n.value, 2), np.int32), np.zeros(n.value + 1, np.int64), np.zeros((m.value, 2), np.int32)
        lib().oracle_pairs(self.h, C.byref(n), C.byref(m), pr.ctypes.data_as(C.POINTER(C.c_int32)),
                           ptr.ctypes.data_as(C.POINTER(C.c_int64)), inc.ctypes.data_as(C.POINTER(C.c_int32)))
        return pr, ptr, inc

    def reduced_system(self, mu, literal=False):
        n = self.n_reduced
        S, rhs = np.zeros((n, n)), np.zeros(n)
        rc = lib().oracle_reduced_system(self.h, mu, int(literal), dptr(S), dptr(rhs))
        return rc, S, rhs

    def solve(self, mu, literal=False):
        delta = np.zeros(self.x_size)
        rc = lib().oracle_solve(self.h, mu, int(literal), dptr(delta))
        return rc, delta

    def max_diagonal(self):
        return lib().oracle_max_diagonal(self.h)

    def gradient_inf_norm(self):
        return lib().oracle_gradient_inf_norm(self.h)

    def gradient_dot(self, delta):
        return lib().oracle_gradient_dot(self.h, dptr(np.ascontiguousarray(delta, np.float64)))

    def mahalanobis(self, x, y_obs=None):
        y_obs = self.scene.y_obs if y_obs is None else y_obs
        return lib().oracle_mahalanobis(self.h, dptr(np.ascontiguousarray(x, np.float64)),
                                        dptr(np.ascontiguousarray(y_obs, np.float64)),
                                        dptr(self.d.kb), dptr(self.d.cw))

    def lm(self, x0=None, max_iterations=20, tau=1e-3, eps1=0.0, eps2=0.0, literal=False):
        s = self.scene
        x0 = s.x0 if x0 is None else x0
        opt = hsba_lm_options(max_iterations=max_iterations, tau=tau, eps1=eps1, eps2=eps2)
        x = np.zeros(self.x_size)
        st = hsba_lm_stats()
        trace = (LMTrace * max(1, max_iterations))()
        rc = lib().oracle_lm(self.h, dptr(np.ascontiguousarray(x0, np.float64)),
                             dptr(np.ascontiguousarray(s.y_obs, np.float64)), dptr(self.d.kb),
                             dptr(self.d.cw), C.byref(opt), int(literal), dptr(x), C.byref(st),
                             trace)
        tr = [dict(mu=t.mu, cost=t.cost, cost_new=t.cost_new, rho=t.rho,
                   step_norm=t.step_norm, accepted=bool(t.accepted))
              for t in list(trace)[:st.iterations]]
        return rc, x, st, tr

    def dense_normal(self, x, y_obs=None):
        y_obs = self.scene.y_obs if y_obs is None else y_obs
        nx, ny = self.x_size, self.y_size
        J, N, g = np.zeros((ny, nx)), np.zeros((nx, nx)), np.zeros(nx)
        lib().oracle_dense_normal(self.h, dptr(np.ascontiguousarray(x, np.float64)),
                                  dptr(np.ascontiguousarray(y_obs, np.float64)), dptr(self.d.kb),
                                  dptr(self.d.cw), dptr(J), dptr(N), dptr(g))
        return J, N, g
