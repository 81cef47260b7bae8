"""Seeded random sweep of the join entry points against the oracle.

Each case draws sizes, a key range (negative and extreme keys included), a
plan (radix bits, passes, LDS table size, S tile) and the flags (order,
RESULT16, left outer), then compares the rows with the oracle's hash join
(joins/CPU_JOIN/hash_cpu restated in oracle/cpu_join.c) as sorted multisets.
"""
import numpy as np
import pytest

import oracle
from eating_one_amd import _abi
from common import same_multiset

pytestmark = pytest.mark.gpu

KEY_MIN, KEY_MAX = -(1 << 31), (1 << 31) - 1


def _case(seed):
    rng = np.random.default_rng(seed)
    nR = int(2 ** rng.uniform(0, 17))
    nS = int(2 ** rng.uniform(0, 18))
    kr = int(2 ** rng.uniform(1, 31))
    kr = max(kr, nR * nS // (1 << 24) + 1)  # at most ~16 M pairs
    lo = int(rng.integers(KEY_MIN, KEY_MAX - kr)) if rng.random() < 0.5 else 0
    R = np.stack([rng.integers(lo, lo + kr, nR, dtype=np.int64), np.arange(nR)], 1).astype(np.int32)
    S = np.stack([rng.integers(lo, lo + kr, nS, dtype=np.int64), np.arange(nS) + 7], 1).astype(np.int32)
    if nR and rng.random() < 0.3:  # extreme keys
        R[rng.integers(0, nR, 3), 0] = [KEY_MIN, KEY_MAX, -1]
        if nS:
            S[rng.integers(0, nS, 3), 0] = [KEY_MIN, KEY_MAX, -1]
    passes = int(rng.integers(0, 3))
    bits = {0: 0, 1: int(rng.integers(1, 12)), 2: int(rng.integers(2, 23))}[passes]
    kw = dict(radix_bits=bits if passes else int(rng.choice([0, -1])), passes=passes,
              lds_tuples=int(rng.choice([0, 4096, 8192, 16384])), deterministic=bool(rng.random() < 0.4),
              result16=bool(rng.random() < 0.3), left_outer=bool(rng.random() < 0.3))
    if kw["radix_bits"] <= 0:
        kw["passes"] = 0
    return R, S, kw


def _want(R, S, outer):
    inner = oracle.hash_join(R, S, nparts=1 << 10)
    if not outer:
        return inner
    miss = S[~np.isin(S[:, 0], R[:, 0])]
    return np.concatenate([inner, np.stack([np.full(len(miss), _abi.EO_NULL_PAYLOAD, np.int32), miss[:, 1]], 1)])


@pytest.mark.parametrize("seed", range(40))
def test_random_hash_join(ctx, seed):
    R, S, kw = _case(1000 + seed)
    want = _want(R, S, kw["left_outer"])
    got = ctx.hash_join(R, S, **kw)
    rows = got[:, [0, 2]] if kw["result16"] else got
    assert same_multiset(rows, want), kw
    if kw["result16"] and len(got):
        keys = got[:, 1]
        assert np.array_equal(keys, got[:, 3])


@pytest.mark.parametrize("seed", range(12))
def test_random_nlj_and_index_joins(ctx, seed):
    R, S, kw = _case(5000 + seed)
    R, S = R[:20000], S[:30000]
    outer = kw["left_outer"]
    want = _want(R, S, outer)
    assert same_multiset(ctx.nlj_join(R, S, left_outer=outer, deterministic=kw["deterministic"]), want)
    for ix in (ctx.hidx_build(R), ctx.sidx_build(R), ctx.sidx_build(R, learned=True)):
        try:
            assert same_multiset(ix.join(S, left_outer=outer, deterministic=kw["deterministic"]), want)
        finally:
            ix.free()


@pytest.mark.parametrize("seed", range(8))
def test_random_stream_join(ctx, seed):
    R, S, kw = _case(9000 + seed)
    rng = np.random.default_rng(seed)
    rt = int(rng.integers(max(1, len(R) // 8), len(R) + 1)) if len(R) else 0  # at most 8 R tiles
    sc = int(rng.integers(max(1, len(S) // 8), len(S) + 1)) if len(S) else 1  # at most 8 S chunks
    outer = kw["left_outer"] and (rt == 0 or rt >= len(R))
    want = _want(R, S, outer)
    for op in ("hash", "nlj"):
        if op == "nlj" and len(R) * len(S) > 2e9:
            continue
        got = ctx.stream_join(R, S, op=op, r_tile=rt, s_chunk=sc, left_outer=outer)
        assert same_multiset(got, want), (op, rt, sc)


def _sampled_case(seed):
    """Relations of 2^20 to 2^21.5 tuples (both sampled), any two-pass split,
    random key range and skew: the sampled partitions' path."""
    rng = np.random.default_rng(seed)
    nR = int(2 ** rng.uniform(20, 21.5))
    nS = int(2 ** rng.uniform(20, 21.5))
    kr = int(2 ** rng.uniform(12, 31))
    kr = max(kr, nR * nS // (1 << 24) + 1)  # at most ~16 M pairs
    lo = int(rng.integers(KEY_MIN, KEY_MAX - kr)) if rng.random() < 0.5 else 0
    R = np.stack([rng.integers(lo, lo + kr, nR, dtype=np.int64), np.arange(nR)], 1).astype(np.int32)
    if rng.random() < 0.5:  # S keys skewed towards a few hot values (few R duplicates: bounded pairs)
        kr = max(kr, nR // 4)
        lo = min(lo, KEY_MAX - kr)
        R[:, 0] = rng.integers(lo, lo + kr, nR, dtype=np.int64).astype(np.int32)
        hot = rng.integers(lo, lo + kr, 16, dtype=np.int64)
        pick = np.where(rng.random(nS) < 0.3, hot[rng.integers(0, 16, nS)], rng.integers(lo, lo + kr, nS, dtype=np.int64))
    else:
        pick = rng.integers(lo, lo + kr, nS, dtype=np.int64)
    S = np.stack([pick, np.arange(nS) + 7], 1).astype(np.int32)
    bits = int(rng.integers(2, 15))
    return R, S, dict(radix_bits=bits, passes=2, lds_tuples=int(rng.choice([0, 8192, 16384])))


@pytest.mark.parametrize("seed", range(8))
def test_random_sampled_partitions(ctx, seed):
    import eating_one_amd as eo

    R, S, kw = _sampled_case(5000 + seed)
    st = eo.JoinStats()
    got = ctx.hash_join(R, S, stats=st, **kw)
    assert st.sampled in (1, 2), kw  # sampled, or sampled and redone exactly
    assert same_multiset(got, oracle.hash_join(R, S, nparts=1 << 10)), kw
