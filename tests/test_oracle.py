"""Pins the CPU oracle (the checker) on the CPU, without a GPU.

- the reference's own data fixtures (joins/join_ori R, S, hash-part-*), regenerated
  from glibc rand() and compared byte-exactly by digest;
- closed forms of the workload shapes (C1 exact-match, C3 histogram product);
- independent reductions (sort + equal_range; numpy cumsum for the scan);
- cross-checks between the four restated CPU_JOIN algorithms;
- planted mistakes in oracle output must be caught by the comparisons.
"""
import numpy as np
import pytest

import oracle
from eating_one_amd import gen
from common import (brute_join, eo_to_ref, join_count_closed_form, join_ori_doc, libc_rand_table, ref_to_eo,
                    same_multiset, sha256, sort_pairs)


@pytest.fixture(scope="module")
def join_ori():
    doc = join_ori_doc()
    R = libc_rand_table(doc["tables"]["R"]["rows"])
    S = libc_rand_table(doc["tables"]["S"]["rows"])
    return doc, R, S


def test_join_ori_tables_regenerate_exactly(join_ori):
    doc, R, S = join_ori
    assert sha256(R) == doc["tables"]["R"]["sha256"]
    assert sha256(S) == doc["tables"]["S"]["sha256"]


@pytest.mark.parametrize("table", ["R", "S"])
def test_oracle_partition_matches_hash_part_files(join_ori, table):
    doc, R, S = join_ori
    T = ref_to_eo(R if table == "R" else S)
    out, start = oracle.partition(T, 4)
    for p in range(4):
        part = eo_to_ref(out[start[p]:start[p + 1]])
        want = doc["partitions"][f"{table}-{p}"]
        assert len(part) == want["rows"]
        assert sha256(part) == want["sha256"], f"partition {table}-{p} differs from hash-part-{table}-{p}"


def test_oracle_partition_bits_equals_modulo(join_ori):
    _, _, S = join_ori
    T = ref_to_eo(S)
    a, sa = oracle.partition(T, 64)
    b, sb = oracle.partition_bits(T, 6, 0)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)


@pytest.mark.parametrize("algo", ["hash", "hidx", "sidx", "nlj", "mt", "sidx_mt"])
def test_join_ori_join_counts(join_ori, algo):
    doc, R, S = join_ori
    Re, Se = ref_to_eo(R), ref_to_eo(S)
    fn = {
        "hash": lambda: oracle.hash_join(Re, Se, nparts=4),
        "hidx": lambda: oracle.hidx_join(Re, Se, nbuckets=1 << 10),
        "sidx": lambda: oracle.sidx_join(Re, Se),
        "nlj": lambda: oracle.nlj_join(Re, Se),
        "mt": lambda: oracle.hash_join_mt(Re, Se, nparts=4, threads=4),
        "sidx_mt": lambda: oracle.sidx_join_mt(Re, Se, threads=4),
    }[algo]
    pairs = fn()
    assert len(pairs) == doc["join"]["pairs"]
    # sum of R.val over all matches, what joins/join_ori/bnlj.cc:73 accumulates
    rval = Re[pairs[:, 0], 0].astype(np.int64)
    assert int(rval.sum()) == doc["join"]["sum_r_val"]
    assert np.array_equal(Re[pairs[:, 0], 0], Se[pairs[:, 1], 0])


def test_oracles_agree_with_sort_reduction_random():
    rng = np.random.default_rng(7)
    for trial in range(20):
        nR = int(rng.integers(0, 3000))
        nS = int(rng.integers(0, 3000))
        kr = int(rng.integers(1, 5000))
        lo = int(rng.integers(-(1 << 31), (1 << 31) - kr)) if trial % 3 == 0 else 0
        R, S = gen.uniform(trial, nR, nS, kr, lo=lo)
        want = brute_join(R, S)
        assert len(want) == join_count_closed_form(R, S) if nR and nS else len(want) == 0
        for got in (oracle.hash_join(R, S, nparts=64), oracle.hidx_join(R, S, nbuckets=1024),
                    oracle.sidx_join(R, S), oracle.nlj_join(R, S), oracle.hash_join_mt(R, S, nparts=64, threads=3),
                    oracle.sidx_join_mt(R, S, threads=3)):
            assert same_multiset(got, want)


def test_oracle_edge_cases():
    e = np.empty((0, 2), dtype=np.int32)
    one = np.array([[5, 0]], dtype=np.int32)
    for fn in (oracle.hash_join, oracle.hidx_join, oracle.sidx_join, oracle.nlj_join):
        assert len(fn(e, e)) == 0
        assert len(fn(one, e)) == 0
        assert len(fn(e, one)) == 0
        assert same_multiset(fn(one, one), np.array([[0, 0]]))
    # all keys equal: full cross product, and the sorted index's last-entry expansion
    R = np.stack([np.full(50, 3), np.arange(50)], 1).astype(np.int32)
    S = np.stack([np.full(40, 3), np.arange(40)], 1).astype(np.int32)
    for fn in (oracle.hash_join, oracle.hidx_join, oracle.sidx_join, oracle.nlj_join):
        assert len(fn(R, S)) == 2000
    # extreme keys
    R = np.array([[-(1 << 31), 0], [(1 << 31) - 1, 1], [0, 2], [-1, 3]], dtype=np.int32)
    S = np.array([[(1 << 31) - 1, 0], [-(1 << 31), 1], [-1, 2], [7, 3]], dtype=np.int32)
    want = np.array([[1, 0], [0, 1], [3, 2]])
    for fn in (oracle.sidx_join, oracle.nlj_join, lambda a, b: oracle.hash_join(a, b, nparts=16),
               lambda a, b: oracle.hidx_join(a, b, nbuckets=16)):
        assert same_multiset(fn(R, S), want)


def test_c1_closed_form():
    R, S = gen.c1()
    pairs = oracle.hash_join(R, S)  # NB_BKTENT = 2^20 partitions, as CPU_JOIN
    assert len(pairs) == 1 << 20
    # every S payload exactly once, paired with the R row holding its key
    assert np.array_equal(np.sort(pairs[:, 1]), np.arange(1 << 20))
    inv = np.empty(1 << 16, dtype=np.int32)
    inv[R[:, 0]] = R[:, 1]
    assert np.array_equal(pairs[:, 0], inv[S[pairs[:, 1], 0]])


def test_c3_shape_closed_form_and_cross_check():
    R, S = gen.c3(nR=1 << 14, nS=1 << 14, key_range=1 << 12)
    n = join_count_closed_form(R, S)
    h = oracle.hash_join(R, S)
    assert len(h) == n
    assert same_multiset(h, oracle.sidx_join(R, S))
    assert same_multiset(h, oracle.hidx_join(R, S, nbuckets=1 << 12))


def test_sidx_build_sorted_permutation():
    R, _ = gen.c3(nR=5000, nS=1, key_range=300)
    idx = oracle.sidx_build(R)
    assert np.all(np.diff(idx[:, 0]) >= 0)
    assert np.array_equal(np.sort(idx[:, 1]), np.arange(5000))
    assert np.array_equal(R[idx[:, 1], 0], idx[:, 0])


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
def test_scan_closed_form(dtype):
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 1000, 4097):
        a = rng.integers(0, 1000, size=n).astype(dtype)
        out, tot = oracle.exclusive_scan(a)
        want = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]).astype(dtype) if n else a
        assert np.array_equal(out, want)
        assert tot == int(a.astype(np.uint64).sum())


def test_planted_mistakes_are_detected():
    """Mistakes planted into oracle output must fail the multiset comparison."""
    R, S = gen.c3(nR=4000, nS=4000, key_range=500)
    good = oracle.hash_join(R, S, nparts=64)
    assert same_multiset(good, brute_join(R, S))
    dropped = good[1:]
    swapped = good[:, ::-1]
    shifted = good.copy()
    shifted[0, 1] += 1
    duplicated = np.concatenate([good[1:], good[:1], good[:1]])
    for bad in (dropped, swapped, shifted, duplicated):
        assert not same_multiset(bad, good)


def test_planted_partition_mistake_is_detected(join_ori):
    doc, R, _ = join_ori
    T = ref_to_eo(R)
    out, start = oracle.partition(T, 4)
    unstable = out.copy()
    a, b = int(start[0]), int(start[0]) + 1
    unstable[[a, b]] = unstable[[b, a]]  # breaks stability only
    part0 = eo_to_ref(unstable[start[0]:start[1]])
    assert sha256(part0) != doc["partitions"]["R-0"]["sha256"]


def test_zipf_sampler_frequencies():
    rng = np.random.default_rng(11)
    n, theta = 1000, 0.9
    r = gen.zipf_ranks(rng, n, theta, 400000)
    assert r.min() >= 0 and r.max() < n
    p = 1.0 / np.arange(1, n + 1) ** theta
    p /= p.sum()
    freq = np.bincount(r, minlength=n) / len(r)
    assert abs(freq[0] - p[0]) < 0.005
    assert abs(freq[:10].sum() - p[:10].sum()) < 0.01
    assert np.all(np.abs(np.cumsum(freq) - np.cumsum(p)) < 0.01)


def test_c2_generator_shape():
    R, S = gen.c2(nR=1 << 14, nS=1 << 15, theta=0.5, selectivity=0.1)
    assert np.array_equal(np.sort(R[:, 0]), np.arange(1 << 14))
    inside = S[:, 0] < (1 << 14)
    assert 0.07 < inside.mean() < 0.13
    assert np.all(S[~inside, 0] >= (1 << 14))
    assert len(oracle.hash_join(R, S, nparts=256)) == int(inside.sum())


def _points(rng, n, lo=0.0, hi=6.0):
    P = np.zeros(n, dtype=oracle.POINT3)
    P["id"] = np.arange(n)
    P["v"] = rng.uniform(lo, hi, size=(n, 3))
    return P


def test_dist_join_oracle_against_numpy_and_boundaries():
    """The distance predicate: numpy's elementwise float64 arithmetic in the same
    order (t0*t0 + t1*t1 + t2*t2, each op rounded) decides every pair alike."""
    rng = np.random.default_rng(3)
    R, S = _points(rng, 300), _points(rng, 400)
    got = oracle.nlj_dist_join(R, S, 1.0)
    t = R["v"][None, :, :] - S["v"][:, None, :]  # r - s
    d = t[..., 0] * t[..., 0] + t[..., 1] * t[..., 1] + t[..., 2] * t[..., 2]
    si, ri = np.nonzero(d < 1.0)
    want = np.stack([R["id"][ri], S["id"][si]], 1)
    assert same_multiset(got, want) and len(got) > 0
    # exact distance 1 is excluded (strict <), one ulp inside is included
    P = np.zeros(3, dtype=oracle.POINT3)
    P["id"] = [0, 1, 2]
    P["v"] = [[0, 0, 0], [1, 0, 0], [np.nextafter(1.0, 0.0), 0, 0]]
    pairs = {tuple(p) for p in oracle.nlj_dist_join(P, P, 1.0).tolist()}
    assert (1, 0) not in pairs and (0, 1) not in pairs
    assert (2, 0) in pairs and (0, 2) in pairs
    # the reference's own data shape: integer values repeated in all 3 coordinates
    # match only when equal
    Q = np.zeros(50, dtype=oracle.POINT3)
    Q["id"] = np.arange(50)
    Q["v"] = np.repeat(rng.integers(0, 20, 50)[:, None], 3, 1)
    n = oracle.nlj_dist_join(Q, Q, 1.0, count_only=True)
    vals = Q["v"][:, 0]
    assert n == int((vals[:, None] == vals[None, :]).sum())
