"""ctypes binding of the C ABI in include/hh.h (libhh.so).

This is the binding a non-pybind caller would write; the tests call the
library through it.  It raises RuntimeError with hh_last_error() on any
non-zero status, and loads libhh.so from this package only: a missing
library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhh.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "hh.h")

PREC = {"fp32": 0, "bf16": 1, "fp16": 2, "fp64": 3}

_lib = None
c_i64, c_vp, c_int, c_dbl = ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -m hignn_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        L.hh_last_error.restype = ctypes.c_char_p
        sig = {
            "hh_init": [], "hh_finalize": [],
            "hh_model_create": [c_vp, c_i64, c_int, c_vp],
            "hh_model_create_instances": [c_vp, c_vp, c_i64, c_int, c_vp],
            "hh_model_load_two_body": [c_vp, c_vp, c_i64],
            "hh_model_set_epsilon": [c_vp, c_dbl], "hh_model_set_eta": [c_vp, c_dbl],
            "hh_model_set_max_iter": [c_vp, c_int],
            "hh_model_set_mat_pool_size_factor": [c_vp, c_int],
            "hh_model_set_max_far_dot_work_node_size": [c_vp, c_int],
            "hh_model_set_max_relative_coord": [c_vp, c_i64],
            "hh_model_set_post_check_flag": [c_vp, c_int],
            "hh_model_set_use_symmetry_flag": [c_vp, c_int],
            "hh_model_set_max_far_field_distance": [c_vp, c_dbl],
            "hh_model_set_precision": [c_vp, c_int],
            "hh_model_update_coord": [c_vp, c_vp, c_i64],
            "hh_model_set_tree_reuse": [c_vp, c_int],
            "hh_model_set_tree_build": [c_vp, c_int],
            "hh_model_dot": [c_vp, c_vp, c_vp, c_i64],
            "hh_model_dense_dot": [c_vp, c_vp, c_vp, c_i64],
            "hh_model_dot_device": [c_vp, c_vp, c_vp, c_i64, c_vp],
            "hh_model_update_coord_device": [c_vp, c_vp, c_i64, c_vp],
            "hh_model_dot_capturable": [c_vp, c_vp, c_vp],
            "hh_model_dot_enqueue": [c_vp, c_vp, c_vp, c_i64, c_vp],
            "hh_model_tune_far_field": [c_vp, c_vp, c_i64, c_vp, c_vp, c_int, c_vp],
            "hh_model_tune_near": [c_vp, c_vp, c_i64, c_vp, c_vp, c_int, c_vp],
            "hh_model_set_near_shape": [c_vp, c_int],
            "hh_model_set_near_workgroups": [c_vp, c_int],
            "hh_model_set_near_overlap": [c_vp, c_int],
            "hh_surrogate_create": [ctypes.c_float, ctypes.c_float, c_vp],
            "hh_surrogate_range": [c_vp, c_vp, c_vp],
            "hh_surrogate_set_params": [c_vp, c_vp],
            "hh_surrogate_get_params": [c_vp, c_vp],
            "hh_surrogate_eval_device": [c_vp, c_vp, c_i64, c_vp, c_vp],
            "hh_surrogate_gradient_device": [c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
            "hh_surrogate_gradient_noalloc": [c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
            "hh_surrogate_train_device": [c_vp, c_vp, c_vp, c_i64, c_int, c_i64, ctypes.c_float,
                                          ctypes.c_float, c_vp, c_vp],
            "hh_model_surrogate_samples": [c_vp, c_i64, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_uint64, c_vp, c_vp, c_vp],
            "hh_model_set_surrogate": [c_vp, c_vp, c_dbl, c_int],
            "hh_model_surrogate_stats": [c_vp, c_vp],
            "hh_model_surrogate_errors": [c_vp, c_vp],
            "hh_device_alloc": [ctypes.c_size_t, c_vp],
            "hh_device_free": [c_vp],
            "hh_memcpy_to_device": [c_vp, c_vp, ctypes.c_size_t],
            "hh_memcpy_to_host": [c_vp, c_vp, ctypes.c_size_t],
            "hh_device_synchronize": [],
            "hh_device_copy": [c_vp, c_vp, ctypes.c_size_t],
            "hh_model_tree_counts": [c_vp, c_vp],
            "hh_model_tree_export": [c_vp] * 7,
            "hh_model_far_ranks": [c_vp, c_vp],
            "hh_model_check_fp64": [c_vp, c_vp, c_vp, c_i64, c_int, c_vp],
            "hh_model_eval_pair": [c_vp, c_vp, c_i64, c_vp, c_int],
            "hh_model_last_timings": [c_vp, c_vp],
            "hh_model_last_profile": [c_vp, c_vp, c_vp, c_vp],
            "hh_tree_build": [c_vp, c_i64, c_int, c_int, c_i64, c_vp],
            "hh_tree_build_device": [c_vp, c_vp, c_i64, c_int, c_int, c_i64, c_vp, c_vp],
            "hh_tree_counts": [c_vp, c_vp],
            "hh_tree_export": [c_vp] * 7,
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int
        L.hh_model_destroy.argtypes = [c_vp]
        L.hh_model_destroy.restype = None
        L.hh_tree_free.argtypes = [c_vp]
        L.hh_tree_free.restype = None
        L.hh_surrogate_destroy.argtypes = [c_vp]
        L.hh_surrogate_destroy.restype = None
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"libhh error {rc}: {lib().hh_last_error().decode()}")


def _p(a):
    return a.ctypes.data_as(c_vp) if a is not None else None


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, 3))


def _export(counts_fn, export_fn, handle, n):
    c = np.zeros(4, dtype=np.int64)
    _chk(counts_fn(handle, _p(c)))
    T, nc, nf, nl = (int(v) for v in c)
    out = dict(tree=np.empty((T, 5), np.uint64), perm=np.empty(n, np.uint64),
               aux=np.empty((T, 6), np.float32), close=np.empty((nc, 2), np.int64),
               far=np.empty((nf, 2), np.int64), leaves=np.empty(nl, np.int64))
    _chk(export_fn(handle, *[_p(out[k]) for k in ("tree", "perm", "aux", "close", "far",
                                                      "leaves")]))
    return out


def tree_build(x, leaf, use_symmetry=True, max_relative_coord=500000):
    """Host-only tree + block lists (no device)."""
    x = _f32(x)
    h = c_vp()
    _chk(lib().hh_tree_build(_p(x), x.shape[0], int(leaf), int(bool(use_symmetry)),
                             int(max_relative_coord), ctypes.byref(h)))
    try:
        return _export(lib().hh_tree_counts, lib().hh_tree_export, h, x.shape[0])
    finally:
        lib().hh_tree_free(h)


def tree_build_device(x, leaf, use_symmetry=True, max_relative_coord=500000, counts=None):
    """Tree + block lists built on the device from positions uploaded to it;
    counts: particles per independent instance (consecutive rows of x)."""
    x = _f32(x)
    counts = np.asarray([x.shape[0]] if counts is None else counts, dtype=np.int64)
    xd = DeviceArray.of(x)
    h = c_vp()
    _chk(lib().hh_tree_build_device(xd.ptr, _p(counts), counts.size, int(leaf),
                                    int(bool(use_symmetry)), int(max_relative_coord), None,
                                    ctypes.byref(h)))
    try:
        return _export(lib().hh_tree_counts, lib().hh_tree_export, h, x.shape[0])
    finally:
        lib().hh_tree_free(h)


class Model:
    """Thin object wrapper over an hh_model handle."""

    def __init__(self, x, leaf, instances=None):
        """instances: particle counts of independent systems whose rows are
        concatenated in x (hh_model_create_instances); None: one system."""
        _chk(lib().hh_init())
        self.x = _f32(x)
        self.n = self.x.shape[0]
        h = c_vp()
        if instances is None:
            _chk(lib().hh_model_create(_p(self.x), self.n, int(leaf), ctypes.byref(h)))
        else:
            counts = np.ascontiguousarray(instances, dtype=np.int64)
            if int(counts.sum()) != self.n:
                raise ValueError("instance counts do not add up to the particle count")
            _chk(lib().hh_model_create_instances(_p(self.x), _p(counts), counts.size, int(leaf),
                                                 ctypes.byref(h)))
        self.h = h
        # held by the object: at interpreter exit the module's globals may be
        # gone before the last model is collected
        self._destroy = lib().hh_model_destroy

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_destroy", None):
            self._destroy(self.h)
            self.h = None

    def load_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
        _chk(lib().hh_model_load_two_body(self.h, _p(w), w.size))

    def set(self, epsilon=None, max_iter=None, use_symmetry=None, max_relative_coord=None,
            precision=None, tree_reuse=None, tree_build=None, near_overlap=None):
        L = lib()
        if near_overlap is not None:
            _chk(L.hh_model_set_near_overlap(self.h, int(bool(near_overlap))))
        if tree_build is not None:
            _chk(L.hh_model_set_tree_build(self.h, {"host": 0, "device": 1}[tree_build]))
        if tree_reuse is not None:
            _chk(L.hh_model_set_tree_reuse(self.h, int(bool(tree_reuse))))
        if epsilon is not None:
            _chk(L.hh_model_set_epsilon(self.h, float(epsilon)))
        if max_iter is not None:
            _chk(L.hh_model_set_max_iter(self.h, int(max_iter)))
        if use_symmetry is not None:
            _chk(L.hh_model_set_use_symmetry_flag(self.h, int(bool(use_symmetry))))
        if max_relative_coord is not None:
            _chk(L.hh_model_set_max_relative_coord(self.h, int(max_relative_coord)))
        if precision is not None:
            _chk(L.hh_model_set_precision(self.h, PREC[precision]))

    def update_coord(self, x=None):
        if x is not None:
            self.x = _f32(x)
        _chk(lib().hh_model_update_coord(self.h, _p(self.x), self.n))

    def dot(self, f):
        f = _f32(f)
        u = np.empty((self.n, 3), np.float32)
        _chk(lib().hh_model_dot(self.h, _p(u), _p(f), self.n))
        return u

    def dense_dot(self, f):
        f = _f32(f)
        u = np.empty((self.n, 3), np.float32)
        _chk(lib().hh_model_dense_dot(self.h, _p(u), _p(f), self.n))
        return u

    def dot_device(self, u_ptr, f_ptr, stream=None):
        _chk(lib().hh_model_dot_device(self.h, c_vp(u_ptr), c_vp(f_ptr), self.n,
                                       c_vp(stream) if stream else None))

    def update_coord_device(self, x_ptr, stream=None):
        _chk(lib().hh_model_update_coord_device(self.h, c_vp(x_ptr), self.n,
                                                c_vp(stream) if stream else None))

    def tune_far_field(self, f_ptr, stream=None):
        """Rows (mode, chunk, far-field ms) of hh_model_tune_far_field; the
        model keeps the fastest."""
        t = np.zeros((8, 3), np.float64)
        rows = ctypes.c_int(0)
        _chk(lib().hh_model_tune_far_field(self.h, c_vp(f_ptr), self.n,
                                           c_vp(stream) if stream else None, _p(t), 8,
                                           ctypes.byref(rows)))
        return [(int(a), int(b), float(c)) for a, b, c in t[:rows.value]]

    def tune_near(self, f_ptr, stream=None):
        """Rows (shape, tiles, threads, workgroups, matvec ms) of
        hh_model_tune_near; the model keeps the fastest 16-bit near-kernel
        tile shape and workgroup count."""
        t = np.zeros((16, 5), np.float64)
        rows = ctypes.c_int(0)
        _chk(lib().hh_model_tune_near(self.h, c_vp(f_ptr), self.n, c_vp(stream) if stream else None,
                                      _p(t), 16, ctypes.byref(rows)))
        return [(int(a), int(b), int(c), int(d), float(e)) for a, b, c, d, e in t[:rows.value]]

    def set_near_shape(self, shape):
        _chk(lib().hh_model_set_near_shape(self.h, int(shape)))

    def set_near_workgroups(self, workgroups):
        """Workgroups of the persistent 16-bit near kernel (0: one per
        compute unit)."""
        _chk(lib().hh_model_set_near_workgroups(self.h, int(workgroups)))

    def dot_capturable(self):
        """(capturable, signature) of hh_model_dot_capturable."""
        cap = ctypes.c_int(0)
        sig = ctypes.c_uint64(0)
        _chk(lib().hh_model_dot_capturable(self.h, ctypes.byref(cap), ctypes.byref(sig)))
        return bool(cap.value), int(sig.value)

    def dot_enqueue(self, u_ptr, f_ptr, stream=None):
        _chk(lib().hh_model_dot_enqueue(self.h, c_vp(u_ptr), c_vp(f_ptr), self.n,
                                        c_vp(stream) if stream else None))

    def export(self):
        return _export(lib().hh_model_tree_counts, lib().hh_model_tree_export, self.h, self.n)

    def counts(self):
        """[tree rows, close pairs, far pairs, leaves] of the last build."""
        c = np.zeros(4, dtype=np.int64)
        _chk(lib().hh_model_tree_counts(self.h, _p(c)))
        return [int(v) for v in c]

    def far_ranks(self):
        c = np.zeros(4, dtype=np.int64)
        _chk(lib().hh_model_tree_counts(self.h, _p(c)))
        r = np.empty(int(c[2]), np.int32)
        _chk(lib().hh_model_far_ranks(self.h, _p(r)))
        return r

    def check_fp64(self, f, targets, convention=0):
        f = _f32(f)
        t = np.ascontiguousarray(targets, dtype=np.int64)
        out = np.empty((t.size, 3), np.float64)
        _chk(lib().hh_model_check_fp64(self.h, _p(f), _p(t), t.size, int(convention), _p(out)))
        return out

    def eval_pair(self, r, precision="fp32"):
        r = _f32(r)
        out = np.empty((r.shape[0], 9), np.float32)
        _chk(lib().hh_model_eval_pair(self.h, _p(r), r.shape[0], _p(out), PREC[precision]))
        return out

    def timings(self):
        t = np.zeros(3, np.float64)
        _chk(lib().hh_model_last_timings(self.h, _p(t)))
        return tuple(float(v) for v in t)

    def surrogate_samples(self, n, r_lo, r_hi, seed=0):
        """(r, targets) DeviceArrays: n displacements, |r| log-uniform in
        [r_lo, r_hi], labelled |r| S(M(r)) by the FP32 MLP."""
        r = DeviceArray((n, 3))
        t = DeviceArray((n, 6))
        _chk(lib().hh_model_surrogate_samples(self.h, int(n), float(r_lo), float(r_hi),
                                              int(seed), r.p, t.p, None))
        return r, t

    def set_surrogate(self, s, tol=2e-3, samples=32):
        self._sur = s  # kept alive while attached
        _chk(lib().hh_model_set_surrogate(self.h, s.h if s is not None else None, float(tol),
                                          int(samples)))

    def surrogate_stats(self):
        """[judged, on surrogate, fell back, surrogate entries, full-MLP far entries]."""
        st = np.zeros(5, np.int64)
        _chk(lib().hh_model_surrogate_stats(self.h, _p(st)))
        return [int(v) for v in st]

    def surrogate_errors(self):
        e = np.empty(int(self.counts()[2]), np.float32)
        _chk(lib().hh_model_surrogate_errors(self.h, _p(e)))
        return e

    def profile(self):
        """(ms, pair evaluations, launches) of [near kernel, ACA passes]."""
        ms = np.zeros(2, np.float64)
        pairs = np.zeros(2, np.int64)
        launches = np.zeros(2, np.int64)
        _chk(lib().hh_model_last_profile(self.h, _p(ms), _p(pairs), _p(launches)))
        return ms, pairs, launches


class DeviceArray:
    """Device memory of the library's own HIP runtime (include/hh_device.h),
    holding a float32 array: callers need no other GPU framework."""

    def __init__(self, shape, data=None, dtype=np.float32):
        self.shape = tuple(int(v) for v in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = c_vp()
        _chk(lib().hh_device_alloc(max(self.nbytes, 1), ctypes.byref(p)))
        self.p = p
        self._free = lib().hh_device_free
        if data is not None:
            self.upload(data)

    @classmethod
    def of(cls, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype=dtype)
        return cls(a.shape, a, dtype)

    @property
    def ptr(self):
        return int(self.p.value or 0)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"{a.nbytes} bytes for a device array of {self.nbytes}")
        _chk(lib().hh_memcpy_to_device(self.p, _p(a), self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _chk(lib().hh_memcpy_to_host(_p(out), self.p, self.nbytes))
        return out

    def __del__(self):
        if getattr(self, "p", None) is not None and getattr(self, "_free", None):
            self._free(self.p)
            self.p = None


def synchronize():
    _chk(lib().hh_device_synchronize())


class Surrogate:
    """The far-field surrogate of include/hh_surrogate.h; r, targets and
    outputs are device pointers (DeviceArray.ptr) of (n,3) and (n,6)
    float32 arrays."""

    NPARAMS = 1382

    def __init__(self, r_lo, r_hi):
        h = c_vp()
        _chk(lib().hh_surrogate_create(float(r_lo), float(r_hi), ctypes.byref(h)))
        self.h = h
        self.r_lo, self.r_hi = float(r_lo), float(r_hi)
        self._destroy = lib().hh_surrogate_destroy

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_destroy", None):
            self._destroy(self.h)
            self.h = None

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
        if p.size != self.NPARAMS:
            raise ValueError(f"{self.NPARAMS} parameters expected")
        _chk(lib().hh_surrogate_set_params(self.h, _p(p)))

    def params(self):
        p = np.empty(self.NPARAMS, np.float32)
        _chk(lib().hh_surrogate_get_params(self.h, _p(p)))
        return p

    def eval(self, r_ptr, n, out_ptr, stream=None):
        _chk(lib().hh_surrogate_eval_device(self.h, c_vp(r_ptr), int(n), c_vp(out_ptr),
                                            c_vp(stream)))

    def gradient(self, r_ptr, t_ptr, n, stream=None):
        g = np.empty(self.NPARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        _chk(lib().hh_surrogate_gradient_device(self.h, c_vp(r_ptr), c_vp(t_ptr), int(n), _p(g),
                                                _p(loss), c_vp(stream)))
        return float(loss[0]), g

    def gradient_noalloc(self, r_ptr, t_ptr, n, grad_ptr, loss_ptr, stream=None):
        """Enqueues the gradient into caller-owned device outputs; allocates
        nothing (the library's fixed workspace holds the block partials)."""
        _chk(lib().hh_surrogate_gradient_noalloc(self.h, c_vp(r_ptr), c_vp(t_ptr), int(n),
                                                 c_vp(grad_ptr), c_vp(loss_ptr), c_vp(stream)))

    def train(self, r_ptr, t_ptr, n, steps, lr, batch=0, lr_final=None, stream=None):
        loss = np.zeros(1, np.float32)
        _chk(lib().hh_surrogate_train_device(self.h, c_vp(r_ptr), c_vp(t_ptr), int(n),
                                             int(steps), int(batch), float(lr),
                                             float(lr if lr_final is None else lr_final),
                                             _p(loss), c_vp(stream)))
        return float(loss[0])


def header_symbols():
    """Function names declared in the headers of include/ (hh.h, hh_dynamics.h)."""
    import glob
    import re
    txt = "".join(open(p).read() for p in sorted(glob.glob(os.path.join(os.path.dirname(HEADER), "*.h"))))
    return sorted(set(re.findall(r"\b(hh_[a-z0-9_]+)\s*\(", txt)))
