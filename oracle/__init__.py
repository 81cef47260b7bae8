"""CPU oracle: a restatement of the reference's CPU_JOIN algorithms in C.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg may import this package, and only as the
checker (or the timed CPU baseline), never as the product path.

Tuples and pairs are numpy arrays of shape (n, 2), dtype int32:
tuple = (key, payload), pair = (R payload, S payload).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P, U64, U32, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        L.oracle_scan_u32.argtypes = [P, P, U64]
        L.oracle_scan_u32.restype = U64
        L.oracle_scan_u64.argtypes = [P, P, U64]
        L.oracle_scan_u64.restype = U64
        L.oracle_partition.argtypes = [P, U64, U32, P, P]
        L.oracle_partition.restype = I
        L.oracle_partition_bits.argtypes = [P, U64, I, I, P, P]
        L.oracle_partition_bits.restype = I
        for name in ("oracle_hash_join", "oracle_hidx_join"):
            getattr(L, name).argtypes = [P, U64, P, U64, U32, P, U64]
            getattr(L, name).restype = U64
        for name in ("oracle_sidx_join", "oracle_nlj_join"):
            getattr(L, name).argtypes = [P, U64, P, U64, P, U64]
            getattr(L, name).restype = U64
        L.oracle_sidx_build.argtypes = [P, U64, P]
        L.oracle_sidx_build.restype = I
        L.oracle_nlj_dist_join.argtypes = [P, U64, P, U64, ctypes.c_double, P, U64]
        L.oracle_nlj_dist_join.restype = U64
        L.oracle_hash_join_mt.argtypes = [P, U64, P, U64, U32, I, P, U64]
        L.oracle_hash_join_mt.restype = U64
        L.oracle_sidx_join_mt.argtypes = [P, U64, P, U64, I, P, U64]
        L.oracle_sidx_join_mt.restype = U64
        _lib = L
    return _lib


def _tuples(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError("tuples must have shape (n, 2)")
    return a


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


FAIL = (1 << 64) - 1


def _run_join(fn, args_before, R, S, extra=(), count_only=False):
    R, S = _tuples(R), _tuples(S)
    n = fn(_ptr(R), len(R), _ptr(S), len(S), *extra, None, 0)
    if n == FAIL:
        raise MemoryError("oracle allocation failed")
    if count_only:
        return n
    out = np.empty((max(n, 1), 2), dtype=np.int32)
    n2 = fn(_ptr(R), len(R), _ptr(S), len(S), *extra, _ptr(out), n)
    assert n2 == n
    return out[:n]


def hash_join(R, S, nparts: int = 1 << 20, count_only: bool = False):
    """joins/CPU_JOIN/hash_cpu/hj.cc:116-228."""
    return _run_join(lib().oracle_hash_join, (), R, S, (nparts,), count_only)


def hidx_join(R, S, nbuckets: int = 1 << 27, count_only: bool = False):
    """joins/CPU_JOIN/inner_cpu/inner-idx-bnlj.cc:142-185."""
    return _run_join(lib().oracle_hidx_join, (), R, S, (nbuckets,), count_only)


def sidx_join(R, S, count_only: bool = False):
    """joins/CPU_JOIN/sortIndex_cpu/inner-idx-bnlj.cc:98-237."""
    return _run_join(lib().oracle_sidx_join, (), R, S, (), count_only)


def nlj_join(R, S, count_only: bool = False):
    """joins/CPU_JOIN/block_cpu/nestLoopJoin.cc:61-77."""
    return _run_join(lib().oracle_nlj_join, (), R, S, (), count_only)


POINT3 = np.dtype([("id", np.int32), ("pad", np.int32), ("v", np.float64, (3,))])  # reference 3d TUPLE, 32 bytes


def nlj_dist_join(R, S, distance: float = 1.0, count_only: bool = False):
    """joins/CPU_JOIN/3d_cpu/nestLoopJoin.cc:73-101: pairs (R id, S id) with
    squared distance < distance^2.  R, S: arrays of POINT3."""
    R = np.ascontiguousarray(R, dtype=POINT3)
    S = np.ascontiguousarray(S, dtype=POINT3)
    fn = lib().oracle_nlj_dist_join
    n = fn(_ptr(R), len(R), _ptr(S), len(S), float(distance), None, 0)
    if count_only:
        return n
    out = np.empty((max(n, 1), 2), dtype=np.int32)
    n2 = fn(_ptr(R), len(R), _ptr(S), len(S), float(distance), _ptr(out), n)
    assert n2 == n
    return out[:n]


def hash_join_mt(R, S, nparts: int = 1 << 20, threads: int = 1, count_only: bool = False):
    """joins/multi_joins/hash/thread-hash-bnlj.cc:146-501 with per-thread output."""
    R, S = _tuples(R), _tuples(S)
    fn = lib().oracle_hash_join_mt
    if count_only:
        n = fn(_ptr(R), len(R), _ptr(S), len(S), nparts, threads, None, 0)
        if n == FAIL:
            raise MemoryError("oracle allocation failed")
        return n
    cap = len(S) * 4 + 1024
    out = np.empty((cap, 2), dtype=np.int32)
    n = fn(_ptr(R), len(R), _ptr(S), len(S), nparts, threads, _ptr(out), cap)
    if n == FAIL:
        raise MemoryError("oracle allocation failed")
    if n > cap:
        out = np.empty((n, 2), dtype=np.int32)
        fn(_ptr(R), len(R), _ptr(S), len(S), nparts, threads, _ptr(out), n)
    return out[:n]


def sidx_join_mt(R, S, threads: int = 1, count_only: bool = False):
    """joins/multi_joins/sortindex/thread-hash-bnlj.cc:180-335 with per-thread output."""
    R, S = _tuples(R), _tuples(S)
    fn = lib().oracle_sidx_join_mt
    n = fn(_ptr(R), len(R), _ptr(S), len(S), threads, None, 0)
    if n == FAIL:
        raise MemoryError("oracle allocation failed")
    if count_only:
        return n
    out = np.empty((max(n, 1), 2), dtype=np.int32)
    if fn(_ptr(R), len(R), _ptr(S), len(S), threads, _ptr(out), n) == FAIL:
        raise MemoryError("oracle allocation failed")
    return out[:n]


def partition(T, nparts: int):
    """Stable partition by key % nparts: joins/CPU_JOIN/hash_cpu/hj.cc:132-181."""
    T = _tuples(T)
    out = np.empty_like(T)
    start = np.empty(nparts + 1, dtype=np.uint64)
    if lib().oracle_partition(_ptr(T), len(T), nparts, _ptr(out), _ptr(start)):
        raise MemoryError
    return out, start


def partition_bits(T, bits: int, shift: int = 0):
    T = _tuples(T)
    out = np.empty_like(T)
    start = np.empty((1 << bits) + 1, dtype=np.uint64)
    if lib().oracle_partition_bits(_ptr(T), len(T), bits, shift, _ptr(out), _ptr(start)):
        raise MemoryError
    return out, start


def sidx_build(R):
    R = _tuples(R)
    out = np.empty_like(R)
    if lib().oracle_sidx_build(_ptr(R), len(R), _ptr(out)):
        raise RuntimeError("sort check failed")
    return out


def exclusive_scan(a):
    a = np.ascontiguousarray(a)
    out = np.empty_like(a)
    if a.dtype == np.uint32:
        tot = lib().oracle_scan_u32(_ptr(a), _ptr(out), len(a))
    elif a.dtype == np.uint64:
        tot = lib().oracle_scan_u64(_ptr(a), _ptr(out), len(a))
    else:
        raise TypeError("uint32 or uint64 expected")
    return out, tot


def sort_pairs(p) -> np.ndarray:
    """Pairs sorted lexicographically: the multiset comparison form."""
    p = np.ascontiguousarray(p, dtype=np.int32).reshape(-1, 2)
    if len(p) == 0:
        return p
    order = np.lexsort((p[:, 1], p[:, 0]))
    return p[order]
