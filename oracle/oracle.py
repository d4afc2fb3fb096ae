"""ctypes front end of the CPU oracle (oracle/hh_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of bench.py, always as the checker or the
timed CPU baseline, never as a product code path.

What pins each oracle function (the reference ships no golden vectors, no
known-answer tests and no fixtures for this path, and its C++ cannot be
built here: Kokkos, Eigen and VTK are absent):

* ``mlp_f32`` / ``mlp_f64`` (python/convert.py:11-20): closed-form special
  cases (r = 0 -> identity; W2 = 0 gives M = (W3 tanh(b2) + b3 + I)/|r|;
  W3 = b3 = 0 gives I/|r|), FP32 vs FP64 agreement, an independent numpy
  evaluation, and -- where the reference's weights are present -- the
  measured properties of the trained net (near-symmetric blocks, Oseen limit).
  libtorch's FP32 rounding order is not pinned (covered by tolerance).
* tree (src/HignnModel.cpp:119-198, src/Build.cpp:17-265): permutation,
  child ranges partition the parent, leaves <= leaf size, every split agrees
  with an independent numpy PCA (left side t <= 0 < right side), BFS
  numbering, dead rows 1,2, brute force on tiny inputs.  The reference's own
  sort tie order and Eigen SVD sign are unpinned (not buildable here) and
  replaced by the documented deterministic choices.
* boxes, admissibility, dual traversal (src/CloseFarCheck.cpp): brute-force
  boxes and admissibility on every listed pair; close and far blocks tile all
  N^2 ordered pairs exactly once.
* close_dot (src/CloseDot.cpp): equals the dense sum over close pairs.
* ACA (src/FarDot.cpp): cross-interpolation (the rank-k approximant is exact
  on the pivot rows and columns), block error vs the dense block of order
  epsilon, symmetric apply equals the transpose.
* dense_dot (src/DenseDot.cpp): equals a numpy evaluation of the same sum.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libhh_oracle.so")
_lib = None

c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(c_vp) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run oracle/build.sh)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_weight_count.restype = c_i64
        L.orc_mlp_f32.argtypes = [c_vp, c_vp, c_i64, c_vp]
        L.orc_mlp_f64.argtypes = [c_vp, c_vp, c_i64, c_vp]
        L.orc_principal_axis.argtypes = [c_vp, c_vp]
        L.orc_model_new.restype = c_vp
        L.orc_model_new.argtypes = [c_vp, c_i64, ctypes.c_int]
        L.orc_model_free.argtypes = [c_vp]
        L.orc_model_set_params.argtypes = [c_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           c_i64, ctypes.c_int]
        L.orc_model_update.argtypes = [c_vp, c_vp]
        L.orc_model_count.restype = c_i64
        L.orc_model_count.argtypes = [c_vp, ctypes.c_int]
        L.orc_model_export.argtypes = [c_vp] * 7
        L.orc_model_is_far.argtypes = [c_vp, c_i64, c_i64]
        L.orc_model_dot.argtypes = [c_vp, c_vp, c_vp, c_vp, ctypes.c_int]
        L.orc_model_far_ranks.argtypes = [c_vp, c_vp]
        L.orc_model_dense_dot.argtypes = [c_vp, c_vp, c_vp, c_vp]
        L.orc_model_dense64_blocks.argtypes = [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]
        L.orc_dense64_ordered.argtypes = [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp]
        L.orc_model_sample_time.restype = ctypes.c_double
        L.orc_model_sample_time.argtypes = [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]
        L.orc_model_aca_factors.restype = ctypes.c_int
        L.orc_model_aca_factors.argtypes = [c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]
        L.orc_model_aca_norms.restype = ctypes.c_int
        L.orc_model_aca_norms.argtypes = [c_vp, c_vp, c_i64, c_vp, c_vp]
        _lib = L
    return _lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def mlp_f32(w, r):
    r = _f32(r, (-1, 3))
    out = np.empty((r.shape[0], 9), dtype=np.float32)
    lib().orc_mlp_f32(_p(_f32(w)), _p(r), r.shape[0], _p(out))
    return out


def mlp_f64(w, r):
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
    out = np.empty((r.shape[0], 9), dtype=np.float64)
    lib().orc_mlp_f64(_p(_f32(w)), _p(r), r.shape[0], _p(out))
    return out


def principal_axis(cov6):
    c = np.ascontiguousarray(cov6, dtype=np.float64)
    u = np.empty(3, dtype=np.float64)
    lib().orc_principal_axis(_p(c), _p(u))
    return u


def dense64_ordered(w, x, f, targets):
    x = _f32(x, (-1, 3))
    f = _f32(f, (-1, 3))
    t = np.ascontiguousarray(targets, dtype=np.int64)
    u = np.empty((t.size, 3), dtype=np.float64)
    lib().orc_dense64_ordered(_p(_f32(w)), _p(x), _p(f), x.shape[0], _p(t), t.size, _p(u))
    return u


class OracleModel:
    """CPU restatement of HignnModel (include/HignnModel.hpp:215-527)."""

    def __init__(self, x, leaf, epsilon=0.05, max_iter=100, use_symmetry=True,
                 max_relative_coord=500000, threads=0):
        self.x = _f32(x, (-1, 3))
        self.n = self.x.shape[0]
        self._free = lib().orc_model_free  # kept: module globals may be gone at exit
        self.h = lib().orc_model_new(_p(self.x), self.n, int(leaf))
        self.params = (float(epsilon), int(max_iter), int(bool(use_symmetry)),
                       int(max_relative_coord), int(threads))
        lib().orc_model_set_params(self.h, *self.params)

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    def update_coord(self, x=None):
        if x is not None:
            self.x = _f32(x, (-1, 3))
        lib().orc_model_update(self.h, _p(self.x))

    def counts(self):
        L = lib()
        return {k: L.orc_model_count(self.h, i)
                for i, k in enumerate(["tree", "close", "far", "leaves"])}

    def export(self):
        c = self.counts()
        tree = np.empty((c["tree"], 5), dtype=np.uint64)
        perm = np.empty(self.n, dtype=np.uint64)
        aux = np.empty((c["tree"], 6), dtype=np.float32)
        close = np.empty((c["close"], 2), dtype=np.int64)
        far = np.empty((c["far"], 2), dtype=np.int64)
        leaves = np.empty(c["leaves"], dtype=np.int64)
        lib().orc_model_export(self.h, _p(tree), _p(perm), _p(aux), _p(close), _p(far),
                               _p(leaves))
        return dict(tree=tree, perm=perm, aux=aux, close=close, far=far, leaves=leaves)

    def is_far(self, a, b):
        return bool(lib().orc_model_is_far(self.h, int(a), int(b)))

    def dot(self, w, f, parts=3):
        u = np.empty((self.n, 3), dtype=np.float64)
        lib().orc_model_dot(self.h, _p(_f32(w)), _p(_f32(f, (-1, 3))), _p(u), int(parts))
        return u

    def far_ranks(self):
        r = np.empty(self.counts()["far"], dtype=np.int32)
        lib().orc_model_far_ranks(self.h, _p(r))
        return r

    def dense_dot(self, w, f):
        u = np.empty((self.n, 3), dtype=np.float64)
        lib().orc_model_dense_dot(self.h, _p(_f32(w)), _p(_f32(f, (-1, 3))), _p(u))
        return u

    def dense64_blocks(self, w, f, targets):
        t = np.ascontiguousarray(targets, dtype=np.int64)
        u = np.empty((t.size, 3), dtype=np.float64)
        lib().orc_model_dense64_blocks(self.h, _p(_f32(w)), _p(_f32(f, (-1, 3))), _p(t),
                                       t.size, _p(u))
        return u

    def sample_time(self, w, f, close_stride, far_stride):
        """Seconds for CloseDot on every close_stride-th close block plus the
        ACA of every far_stride-th far block, and the pair evaluations done
        (near, far)."""
        work = np.zeros(2, dtype=np.int64)
        t = lib().orc_model_sample_time(self.h, _p(_f32(w)), _p(_f32(f, (-1, 3))),
                                        int(close_stride), int(far_stride), _p(work))
        return t, int(work[0]), int(work[1])

    def aca_factors(self, w, b, max_iter):
        e = self.export()
        tree = e["tree"].astype(np.int64)
        I, J = e["far"][b]
        m = tree[I, 3] - tree[I, 2]
        n = tree[J, 3] - tree[J, 2]
        C = np.zeros((max_iter, m, 3, 3))
        Q = np.zeros((max_iter, n, 3, 3))
        ip = np.zeros(max_iter, dtype=np.int64)
        jp = np.zeros(max_iter + 1, dtype=np.int64)
        k = lib().orc_model_aca_factors(self.h, _p(_f32(w)), int(b), _p(C), _p(Q), _p(ip),
                                        _p(jp))
        return dict(k=k, C=C[:k], Q=Q[:k], ipiv=ip[:k], jpiv=jp[:k + 1], I=int(I), J=int(J))

    def aca_norms(self, w, b, max_iter):
        """Per-iteration nu2 and mu2 (after the iteration's update) of far
        block b's ACA."""
        nu2 = np.zeros(max_iter)
        mu2 = np.zeros(max_iter)
        k = lib().orc_model_aca_norms(self.h, _p(_f32(w)), int(b), _p(nu2), _p(mu2))
        return nu2[:k], mu2[:k]


def rel_l2(u, ref):
    u = np.asarray(u, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.sqrt(np.sum((u - ref) ** 2) / np.sum(ref ** 2)))
