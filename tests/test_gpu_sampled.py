"""GPU parity of the hash join's sampled partitions: two-pass partitions whose
regions are sized from a 1/16 sample of the input instead of a histogram pass
(runs reserved with atomics, gaps between partitions), and the exact redo
when the sample underestimated a region.  Every result is compared with the
CPU oracle as a sorted multiset with an identical count.
"""
import numpy as np
import pytest

import eating_one_amd as eo
import oracle
from eating_one_amd import _abi, gen
from common import join_count_closed_form, same_multiset

pytestmark = pytest.mark.gpu

# sizes off the 16384-tuple tile and odd: the last tile is partial and the
# last tuple has no pair partner
NR = (1 << 20) + 777
NS = (1 << 21) + 4097


def _join(ctx, R, S, **kw):
    st = eo.JoinStats()
    got = ctx.hash_join(R, S, stats=st, **kw)
    return got, st


@pytest.mark.parametrize("theta", [0.0, 0.9])
@pytest.mark.parametrize("bits", [2, 9, 13, 14])
def test_sampled_partitions_equal_oracle(ctx, bits, theta):
    R, S = gen.c2(seed=0x5A + bits, nR=NR, nS=NS, theta=theta, selectivity=1.0)
    got, st = _join(ctx, R, S, radix_bits=bits, passes=2)
    assert st.passes == 2 and st.radix_bits == bits
    assert st.sampled == 1  # sized from the sample, no redo
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == join_count_closed_form(R, S)
    assert same_multiset(got, want)


def test_sampled_partitions_build_on_the_smaller_side(ctx):
    """|S| < |R|: the inner join builds on S; both relations still large
    enough for sampled partitions."""
    R, S = gen.c3(seed=0x5B, nR=NS, nS=NR, key_range=1 << 21)
    got, st = _join(ctx, R, S, radix_bits=13, passes=2)
    assert st.sampled == 1
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == join_count_closed_form(R, S)
    assert same_multiset(got, want)


def test_sampled_and_exact_partitions_give_the_same_rows(ctx):
    R, S = gen.c2(seed=0x5C, nR=NR, nS=NS, theta=0.5, selectivity=0.1)
    a, sa = _join(ctx, R, S, radix_bits=14, passes=2)
    b, sb = _join(ctx, R, S, radix_bits=14, passes=2, flags=_abi.EO_F_EXACT_PARTITIONS)
    assert sa.sampled == 1 and sb.sampled == 0
    assert same_multiset(a, b)
    # paths that keep exact partitions: deterministic order, left outer join
    _, sd = _join(ctx, R, S, radix_bits=14, passes=2, deterministic=True)
    _, so = _join(ctx, R, S, radix_bits=14, passes=2, left_outer=True)
    assert sd.sampled == 0 and so.sampled == 0


def _sampled_lines(t, j):
    """Line of tile t sampled for group j (the partition's sampling pattern,
    restated to build an input that defeats it)."""
    h = ((t * 64 + j) * 0x9E3779B1) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x85EBCA77) & 0xFFFFFFFF
    h ^= h >> 13
    return j * 16 + (h >> 28)


@pytest.mark.parametrize("side", ["S", "R"])
def test_underestimated_region_redoes_the_join_exactly(ctx, side):
    """Every line the sample reads holds keys spread over all partitions and
    every other line one hot key: the sample sees almost nothing of the hot
    partition, whose region overflows; the join is redone with exact
    partitions and still returns every row."""
    rng = np.random.default_rng(0x5D)
    n = 1 << 21
    keys = np.full(n, 5, dtype=np.int32)
    tiles = n // 16384
    for t in range(tiles):
        for j in range(64):
            line = _sampled_lines(t, j)
            b = t * 16384 + line * 16
            keys[b:b + 16] = rng.integers(0, 1 << 20, size=16, dtype=np.int32)
    big = np.stack([keys, np.arange(n, dtype=np.int32)], 1)
    small = np.stack([rng.permutation(1 << 20).astype(np.int32), np.arange(1 << 20, dtype=np.int32)], 1)
    R, S = (small, big) if side == "S" else (big, small)
    got, st = _join(ctx, R, S, radix_bits=14, passes=2)
    assert st.sampled == 2  # the sampled attempt overflowed, the exact one ran
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == join_count_closed_form(R, S)
    assert same_multiset(got, want)


@pytest.mark.parametrize("nR,nS", [(300_001, (1 << 22) + 3), ((1 << 21) + 9, 500_003)])
def test_one_side_sampled_the_other_exact(ctx, nR, nS):
    """A relation under 2^20 tuples keeps exact (contiguous) partitions while
    the other one's are sized from the sample: the probe reads each side's
    partitions through its own [start, end) arrays."""
    R, S = gen.c3(seed=nR % 97, nR=nR, nS=nS, key_range=1 << 20)
    got, st = _join(ctx, R, S, radix_bits=14, passes=2)
    assert st.passes == 2 and st.sampled == 1
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == join_count_closed_form(R, S)
    assert same_multiset(got, want)


@pytest.mark.parametrize("n,sampled", [(1 << 20, 1), ((1 << 20) - 1, 0)])
def test_sampled_threshold(ctx, n, sampled):
    """Relations of 2^20 tuples take sampled partitions, one tuple fewer the
    exact ones (both relations at the size, two passes forced)."""
    R, S = gen.c2(seed=n, nR=n, nS=n)
    got, st = _join(ctx, R, S, radix_bits=14, passes=2)
    assert st.sampled == sampled
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == n
    assert same_multiset(got, want)


def test_sampled_single_hot_key(ctx):
    """Every probe key equal: one partition holds the whole of S (every line
    alike, so the sample has no error there), the others are empty; every S
    tuple pairs with the one R tuple of that key."""
    nR, nS = 1 << 20, (1 << 21) + 3
    keys = np.random.default_rng(0x5E).permutation(nR).astype(np.int32)
    R = np.stack([keys, np.arange(nR, dtype=np.int32)], 1)
    S = np.stack([np.full(nS, 777_777, dtype=np.int32), np.arange(nS, dtype=np.int32)], 1)
    got, st = _join(ctx, R, S, radix_bits=14, passes=2)
    assert st.sampled == 1
    assert len(got) == nS
    assert np.array_equal(np.sort(got[:, 1]), np.arange(nS, dtype=np.int32))
    assert (got[:, 0] == int(np.flatnonzero(keys == 777_777)[0])).all()


@pytest.mark.parametrize("layout", ["sorted", "descending", "half_sorted", "sorted_dups", "runs_of_4"])
def test_clumped_input_takes_the_exact_partitions(ctx, layout):
    """Sorted keys put whole 128-byte lines into one high digit: the sample,
    which reads whole lines, sees each digit in steps of 16 tuples.  The
    layout measures that clumping and, finding margins too wide for the
    buffers, hands over to the exact partition before any tuple moves; the
    rows are the oracle's."""
    rng = np.random.default_rng(0x5F)
    nR, nS = (1 << 20) + 5, (1 << 21) + 11
    keys = np.arange(nR, dtype=np.int32)
    if layout == "descending":
        keys = keys[::-1].copy()
    elif layout == "half_sorted":  # the second half of the tiles in random order
        keys[nR // 2:] = rng.permutation(keys[nR // 2:])
    R = np.stack([keys, rng.integers(0, 1 << 30, nR, dtype=np.int32)], 1)
    skeys = rng.integers(0, nR, nS).astype(np.int32)
    if layout == "sorted_dups":  # S grouped by key too: each R key about twice in a row
        skeys.sort()
    elif layout == "runs_of_4":  # R in random order, S keys in runs of four (variance ratio ~4)
        keys[:] = rng.permutation(nR)
        R[:, 0] = keys
        skeys = np.repeat(skeys[: nS // 4 + 1], 4)[:nS]
    S = np.stack([skeys, np.arange(nS, dtype=np.int32)], 1)
    got, st = _join(ctx, R, S, radix_bits=14, passes=2)
    assert st.sampled == 2
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == nS
    assert same_multiset(got, want)


@pytest.mark.parametrize("layout", ["random", "adjacent_pairs"])
def test_unclumped_input_keeps_sampled_partitions(ctx, layout):
    """The same sizes in random order: no clumping, or S keys in adjacent
    pairs (variance ratio ~2, within what the buffers are sized for, so the
    margins widen and no redo follows): sampled partitions."""
    rng = np.random.default_rng(0x60)
    nR, nS = (1 << 20) + 5, (1 << 21) + 11
    R = np.stack([rng.permutation(nR).astype(np.int32), np.arange(nR, dtype=np.int32)], 1)
    skeys = rng.integers(0, nR, nS).astype(np.int32)
    if layout == "adjacent_pairs":
        skeys = np.repeat(skeys[: nS // 2 + 1], 2)[:nS]
    S = np.stack([skeys, np.arange(nS, dtype=np.int32)], 1)
    got, st = _join(ctx, R, S, radix_bits=14, passes=2)
    assert st.sampled == 1
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == nS
    assert same_multiset(got, want)


@pytest.mark.parametrize("layout", ["sorted", "one_key", "runs_of_64"])
@pytest.mark.parametrize("bits,passes", [(6, 1), (11, 1), (14, 2)])
def test_exact_partitions_on_uniform_waves(ctx, layout, bits, passes):
    """Exact partitions of inputs where whole waves hold one digit: sorted
    keys (the two-pass first digit is the key's high part), one probe key
    for every S tuple, S in runs of 64 equal keys.  The histogram and the
    scatter rank such a wave with one atomic; sizes off the tile grid and
    odd, so partial waves and the odd last tuple take part.  Rows equal the
    oracle's."""
    rng = np.random.default_rng(0x61 + bits)
    nR, nS = (1 << 19) + 333, (1 << 20) + 4097
    if layout == "sorted":
        rk = np.arange(nR, dtype=np.int32)
        sk = np.sort(rng.integers(0, nR, nS)).astype(np.int32)
    elif layout == "one_key":
        rk = np.arange(nR, dtype=np.int32)
        sk = np.full(nS, 4242, dtype=np.int32)
    else:  # unique R keys in random order, S in runs of 64 equal keys
        rk = rng.permutation(nR).astype(np.int32)
        sk = np.repeat(rng.integers(0, nR, nS // 64 + 1).astype(np.int32), 64)[:nS]
    R = np.stack([rk, np.arange(nR, dtype=np.int32)], 1)
    S = np.stack([sk, np.arange(nS, dtype=np.int32)], 1)
    got, st = _join(ctx, R, S, radix_bits=bits, passes=passes, flags=_abi.EO_F_EXACT_PARTITIONS)
    assert st.sampled == 0 and st.passes == passes
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == join_count_closed_form(R, S)
    assert same_multiset(got, want)
