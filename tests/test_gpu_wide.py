"""Counts past 32 bits and grids past one tile per block.

The reference counts matches in 32-bit ints (`int resultVal`,
joins/CPU_JOIN/hash_cpu/hj.cc:190) and caps its output at JT_SIZE
(joins/hashJoin_gpu/main_street/tuple.h:39).  Here every join counts in 64
bits: 2^17 x 2^17 tuples on one key give 2^34 matches, counted (out = NULL)
by every operator.
"""
import numpy as np
import pytest

import eating_one_amd as eo
import oracle
from eating_one_amd import DeviceArray, gen
from common import same_multiset

pytestmark = pytest.mark.gpu

N = 1 << 17


def _one_key(n, key=7):
    return np.stack([np.full(n, key, np.int32), np.arange(n, dtype=np.int32)], 1)


def test_counts_above_2_32_every_join(ctx):
    R, S = _one_key(N), _one_key(N)
    want = N * N
    assert want == 1 << 34
    print("hash join, fused and deterministic", flush=True)
    assert ctx.hash_join(R, S, count_only=True) == want
    assert ctx.hash_join(R, S, count_only=True, deterministic=True) == want
    print("nested-loop join, fused and deterministic", flush=True)
    assert ctx.nlj_join(R, S, count_only=True) == want
    assert ctx.nlj_join(R, S, count_only=True, deterministic=True) == want
    print("streaming joins", flush=True)
    assert ctx.stream_join(R, S, op="hash", s_chunk=1 << 15, count_only=True) == want
    assert ctx.stream_join(R, S, op="nlj", s_chunk=1 << 15, count_only=True) == want
    print("hash index", flush=True)
    ix = ctx.hidx_build(R)
    assert ix.join(S, count_only=True) == want
    assert ix.join(S, count_only=True, deterministic=True) == want
    ix.free()
    print("sorted and learned index", flush=True)
    for learned in (False, True):
        ix = ctx.sidx_build(R, learned=learned)
        assert ix.join(S, count_only=True) == want
        assert ix.join(S, count_only=True, binary_search=True) == want
        ix.free()
    print("batch", flush=True)
    dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
    b = ctx.batch()
    res = b.run([eo.request(R=dR, nR=N, S=dS, nS=N, tag=1), eo.request(R=dR, nR=N, S=dS, nS=N // 2, tag=2)])
    assert res == [(1, want, 0), (2, want // 2, 0)]
    b.free()


def test_counts_above_2_32_capacity_report(ctx):
    """A too small output buffer reports the whole 64-bit count with EO_ECAPACITY."""
    R, S = _one_key(N), _one_key(N)
    dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
    dout = DeviceArray(ctx, 8 * 1000)
    n, st = ctx.hash_join_device(dR, N, dS, N, dout, 1000)
    assert st == eo._abi.EO_ECAPACITY and n == 1 << 34
    rows = dout.to_host(np.int32, (1000, 2))
    assert (rows >= 0).all() and (rows < N).all()


def test_fused_probe_more_tiles_than_blocks(ctx):
    """Four partitions with 8192-tuple S spans give 1024 tiles, more than the
    probe's grid (two blocks per CU): blocks run several tiles each."""
    R, S = gen.c2(seed=71, nR=60000, nS=1 << 23, theta=0.5)
    st = eo._abi.JoinStats()
    got = ctx.hash_join(R, S, radix_bits=2, passes=1, s_tile=8192, stats=st)
    assert st.tiles >= 1024
    want = oracle.hash_join(R, S, nparts=1 << 10)
    assert len(got) == len(want)
    assert same_multiset(got, want)


def test_many_partitions_few_tuples(ctx):
    """2^22 partitions for 60000 tuples: the probe's grid follows the tuples, not the partitions."""
    R, S = gen.c3(seed=72, nR=60000, nS=90000, key_range=1 << 20)
    for det in (False, True):
        got = ctx.hash_join(R, S, radix_bits=22, passes=2, deterministic=det)
        assert same_multiset(got, oracle.hash_join(R, S, nparts=1 << 12))
