"""Shared test helpers: fixture regeneration and multiset comparison."""
import ctypes
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def join_ori_doc():
    return json.load(open(os.path.join(GOLDEN, "join_ori.json")))


def libc_rand_table(n: int) -> np.ndarray:
    """[(i, rand() % 100) for i in range(n)] after glibc srand(1): the generator
    of the reference's joins/join_ori/R and S tables (reference layout {id, val})."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    vals = np.fromiter((libc.rand() % 100 for _ in range(n)), dtype=np.int32, count=n)
    out = np.empty((n, 2), dtype=np.int32)
    out[:, 0] = np.arange(n, dtype=np.int32)
    out[:, 1] = vals
    return out


def ref_to_eo(t: np.ndarray) -> np.ndarray:
    """Reference TUPLE {key = id, val} -> eo tuple (key = val, payload = id)."""
    return np.ascontiguousarray(t[:, ::-1])


def eo_to_ref(t: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(t[:, ::-1])


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sort_pairs(p: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.int32).reshape(-1, 2)
    if len(p) == 0:
        return p
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def same_multiset(a: np.ndarray, b: np.ndarray) -> bool:
    a = sort_pairs(a)
    b = sort_pairs(b)
    return a.shape == b.shape and bool(np.array_equal(a, b))


def brute_join(R: np.ndarray, S: np.ndarray) -> np.ndarray:
    """Equi-join by sort + equal_range (std::sort/std::equal_range reduction), independent of the oracle."""
    R = np.asarray(R)
    S = np.asarray(S)
    if len(R) == 0 or len(S) == 0:
        return np.empty((0, 2), dtype=np.int32)
    order = np.argsort(R[:, 0], kind="stable")
    rk = R[order, 0]
    lo = np.searchsorted(rk, S[:, 0], side="left")
    hi = np.searchsorted(rk, S[:, 0], side="right")
    cnt = hi - lo
    total = int(cnt.sum())
    out = np.empty((total, 2), dtype=np.int32)
    s_idx = np.repeat(np.arange(len(S)), cnt)
    start = np.repeat(lo, cnt)
    offs = np.arange(total) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    out[:, 0] = R[order[start + offs], 1]
    out[:, 1] = S[s_idx, 1]
    return out


def join_count_closed_form(R: np.ndarray, S: np.ndarray) -> int:
    """|R join S| = sum_k cR[k] * cS[k] from the two key histograms."""
    rk, rc = np.unique(np.asarray(R)[:, 0], return_counts=True)
    sk, sc = np.unique(np.asarray(S)[:, 0], return_counts=True)
    _, ri, si = np.intersect1d(rk, sk, assume_unique=True, return_indices=True)
    return int((rc[ri].astype(np.int64) * sc[si].astype(np.int64)).sum())
