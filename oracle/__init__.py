"""CPU oracle for the exact kNN join -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, and only as the checker or the timed CPU baseline;
the library under ``knn_sweet_amd/`` never calls it.

``bruteforce`` is the north-star oracle: binary64 brute force with the
(squared distance, index) key defined in ``knnj_oracle.c``.  ``ti_join`` is
the CPU triangle-inequality join (a restatement of the reference algorithm,
src/knnjoin.cu:166-525 and 762-858, with its exactness defects fixed).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    i64, i32 = ctypes.c_int64, ctypes.c_int32
    lib.knnj_oracle_bruteforce.restype = ctypes.c_int
    lib.knnj_oracle_bruteforce.argtypes = [
        _f32p, i64, i64, _f32p, i64, i64, i32, i32, _i32p, _f32p, _f64p, i32]
    lib.knnj_oracle_certify.restype = ctypes.c_int64
    lib.knnj_oracle_certify.argtypes = [_f32p, i64, _f32p, i64, i64, i32, i32, _i32p]
    lib.knnj_oracle_pair_sq64.restype = None
    lib.knnj_oracle_pair_sq64.argtypes = [_f32p, i64, _f32p, i64, i32, _i32p, _i32p, i64, _f64p]
    lib.knnj_oracle_ti_join.restype = ctypes.c_int
    lib.knnj_oracle_ti_join.argtypes = [
        _f32p, i64, i64, _f32p, i64, i64, i32, i32, i32, i32, i32, ctypes.c_uint64,
        _i32p, _f32p, _f64p, ctypes.POINTER(ctypes.c_int64), i32, _f64p]
    _lib = lib
    return lib


def _as_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.ndim != 2:
        raise ValueError("expected a 2-D array")
    return x


def bruteforce(R, Q, k: int, nthreads: int = 0):
    """FP64 brute-force exact kNN: returns (idx int32 [nQ,k], dist f32 [nQ,k], sq f64 [nQ,k])."""
    lib = _load()
    R, Q = _as_f32(R), _as_f32(Q)
    nR, d = R.shape
    nQ = Q.shape[0]
    if Q.shape[1] != d:
        raise ValueError("R and Q differ in dimension")
    idx = np.empty((nQ, k), np.int32)
    dist = np.empty((nQ, k), np.float32)
    sq = np.empty((nQ, k), np.float64)
    rc = lib.knnj_oracle_bruteforce(R, nR, d, Q, nQ, d, d, k, idx, dist, sq, nthreads)
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    return idx, dist, sq


def certify(R, Q, idx) -> int:
    """Adjacent pairs whose exact (binary128) order differs from the row order."""
    lib = _load()
    R, Q = _as_f32(R), _as_f32(Q)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    return int(lib.knnj_oracle_certify(R, R.shape[1], Q, Q.shape[0], Q.shape[1], Q.shape[1],
                                       idx.shape[1], idx))


def pair_sq64(R, Q, qi, ri):
    """The oracle's binary64 squared distance for explicit (query, reference) pairs."""
    lib = _load()
    R, Q = _as_f32(R), _as_f32(Q)
    qi = np.ascontiguousarray(qi, dtype=np.int32).ravel()
    ri = np.ascontiguousarray(ri, dtype=np.int32).ravel()
    out = np.empty(qi.shape[0], np.float64)
    lib.knnj_oracle_pair_sq64(R, R.shape[1], Q, Q.shape[1], R.shape[1], qi, ri, qi.shape[0], out)
    return out


def ti_join(R, Q, k: int, m_q: int = 0, m_r: int = 0, trials: int = 10, seed: int = 2015,
            nthreads: int = 0, phases: bool = False):
    """CPU triangle-inequality join: returns (idx, dist, sq, n_distance_evaluations),
    plus the seconds of (index build and bounds, per-query scan) with phases=True."""
    lib = _load()
    R, Q = _as_f32(R), _as_f32(Q)
    nR, d = R.shape
    nQ = Q.shape[0]
    m_q = m_q or max(1, int(2 * np.sqrt(max(nQ, 1))))
    m_r = m_r or max(1, int(2 * np.sqrt(nR)))
    idx = np.empty((nQ, k), np.int32)
    dist = np.empty((nQ, k), np.float32)
    sq = np.empty((nQ, k), np.float64)
    ev = ctypes.c_int64(0)
    tp = np.zeros(2, np.float64)
    rc = lib.knnj_oracle_ti_join(R, nR, d, Q, nQ, d, d, k, m_q, m_r, trials, seed, idx, dist, sq,
                                 ctypes.byref(ev), nthreads, tp)
    if rc != 0:
        raise ValueError("ti_join rejected the arguments")
    if phases:
        return idx, dist, sq, int(ev.value), (float(tp[0]), float(tp[1]))
    return idx, dist, sq, int(ev.value)
