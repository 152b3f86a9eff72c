"""ctypes view of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, by ``__graft_entry__.smoke()`` and by
the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package.  See
``mcr_oracle.h`` for what each function restates and how it is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB_PATH selects the sanitizer build (tests/sanitize.sh)
LIB_PATH = os.environ.get("ORACLE_LIB_PATH") or os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class ModelF32(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("L", _f32p), ("mu", _f32p), ("v", _f32p),
                ("mp", ctypes.c_float), ("thr", ctypes.c_float)]


class ModelF64(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("L", _f64p), ("mu", _f64p), ("v", _f64p),
                ("mp", ctypes.c_double), ("x", ctypes.c_double)]


def _declare(L: ctypes.CDLL) -> None:
    L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.orc_neg_log_u24.argtypes = [ctypes.c_uint32]
    L.orc_neg_log_u24.restype = ctypes.c_float
    L.orc_sqrt_twice.argtypes = [ctypes.c_float]
    L.orc_sqrt_twice.restype = ctypes.c_float
    L.orc_sincos_turn.argtypes = [ctypes.c_uint32, _f32p, _f32p]
    L.orc_box_muller_f32.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p]
    L.orc_normals_f32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _f32p]
    L.orc_normals_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _f64p]
    L.orc_normals_batch_f32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                        ctypes.c_int, _f32p]
    L.orc_cholesky.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                               ctypes.c_double, _f64p, ctypes.POINTER(ctypes.c_int)]
    L.orc_cholesky.restype = ctypes.c_int
    L.orc_factor.argtypes = [ctypes.c_int, _f64p, ctypes.c_int, _f64p, _f64p,
                             ctypes.POINTER(ctypes.c_int)]
    L.orc_factor.restype = ctypes.c_int
    L.orc_simulate_f32.argtypes = [ctypes.POINTER(ModelF32), ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_int64, ctypes.c_int, _i32p, _f32p, _i64p, _i64p,
                                   ctypes.c_int]
    L.orc_simulate_f32.restype = ctypes.c_int64
    L.orc_simulate_f64.argtypes = [ctypes.POINTER(ModelF64), ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_int64, ctypes.c_int, _i32p, _f64p, ctypes.c_int]
    L.orc_simulate_f64.restype = ctypes.c_int64
    L.orc_stats.argtypes = [ctypes.c_int64, ctypes.c_int64, _f64p]
    L.orc_fixed32.argtypes = [ctypes.c_float]
    L.orc_fixed32.restype = ctypes.c_int64


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t) if a is not None else None


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return o


def neg_log_u24(a: int) -> float:
    return lib().orc_neg_log_u24(a)


def sqrt_twice(h: float) -> float:
    return lib().orc_sqrt_twice(h)


def sincos_turn(b: int):
    c, s = ctypes.c_float(), ctypes.c_float()
    lib().orc_sincos_turn(b, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


def normals_f32(seed: int, path0: int, npaths: int, n: int) -> np.ndarray:
    z = np.zeros((npaths, n), dtype=np.float32)
    lib().orc_normals_batch_f32(seed, path0, npaths, n, _p(z, _f32p))
    return z


def normals_f64(seed: int, path: int, n: int) -> np.ndarray:
    z = np.zeros(n, dtype=np.float64)
    lib().orc_normals_f64(seed, path, n, _p(z, _f64p))
    return z


def cholesky(A: np.ndarray, delta: float = 0.0, clamp: bool = False, clamp_value: float = 0.0):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    L = np.zeros((n, n), dtype=np.float64)
    nc = ctypes.c_int(0)
    st = lib().orc_cholesky(n, _p(A, _f64p), n, delta, int(clamp), clamp_value, _p(L, _f64p),
                            ctypes.byref(nc))
    return st, L, nc.value


def factor(A: np.ndarray):
    """Jitter policy; returns (attempts, L, delta, clamped)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    L = np.zeros((n, n), dtype=np.float64)
    d = ctypes.c_double(0.0)
    nc = ctypes.c_int(0)
    att = lib().orc_factor(n, _p(A, _f64p), n, _p(L, _f64p), ctypes.byref(d), ctypes.byref(nc))
    return att, L, d.value, nc.value


class F32Model:
    """fp32 model arrays as the HIP model holds them (L, mu, v, w^T mu, -x)."""

    def __init__(self, L32, mu32, v32, mp32, thr32):
        self.n = int(L32.shape[0])
        self.L = np.ascontiguousarray(L32, dtype=np.float32)
        self.mu = np.ascontiguousarray(mu32, dtype=np.float32)
        self.v = np.ascontiguousarray(v32, dtype=np.float32)
        self.mp = np.float32(mp32)
        self.thr = np.float32(thr32)
        self._s = ModelF32(self.n, _p(self.L, _f32p), _p(self.mu, _f32p), _p(self.v, _f32p),
                           float(self.mp), float(self.thr))

    @classmethod
    def from_f64(cls, L, mu, w, x):
        L = np.asarray(L, dtype=np.float64)
        v = L.T @ np.asarray(w, dtype=np.float64)
        mp = float(np.dot(w, mu))
        return cls(L.astype(np.float32), np.asarray(mu).astype(np.float32), v.astype(np.float32),
                   np.float32(mp), np.float32(-x))


def simulate_f32(model: F32Model, seed: int, path0: int, npaths: int, k: int,
                 want_S=False, want_loss=False, want_hist=False, nthreads: int = 1):
    S = np.zeros(npaths, dtype=np.int32) if want_S else None
    loss = np.zeros(npaths, dtype=np.float32) if want_loss else None
    hist = np.zeros(model.n + 1, dtype=np.int64) if want_hist else None
    near = ctypes.c_int64(0)
    cnt = lib().orc_simulate_f32(ctypes.byref(model._s), seed, path0, npaths, k, _p(S, _i32p),
                                 _p(loss, _f32p), ctypes.byref(near), _p(hist, _i64p), nthreads)
    return {"count": int(cnt), "near": int(near.value), "S": S, "loss": loss, "hist": hist}


def simulate_f64(L, mu, w, x, seed: int, path0: int, npaths: int, k: int, want_S=False,
                 want_loss=False, nthreads: int = 1):
    L = np.ascontiguousarray(L, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    v = np.ascontiguousarray(L.T @ w)
    m = ModelF64(L.shape[0], _p(L, _f64p), _p(mu, _f64p), _p(v, _f64p), float(np.dot(w, mu)), x)
    S = np.zeros(npaths, dtype=np.int32) if want_S else None
    loss = np.zeros(npaths, dtype=np.float64) if want_loss else None
    cnt = lib().orc_simulate_f64(ctypes.byref(m), seed, path0, npaths, k, _p(S, _i32p),
                                 _p(loss, _f64p), nthreads)
    return {"count": int(cnt), "S": S, "loss": loss}


def stats(count: int, M: int):
    o = np.zeros(4)
    lib().orc_stats(count, M, _p(o, _f64p))
    return tuple(float(t) for t in o)


def fixed32(v: float) -> int:
    return int(lib().orc_fixed32(v))
