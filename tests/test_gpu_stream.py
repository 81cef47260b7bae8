"""GPU parity of the out-of-core streaming join (eo_stream_join) through the C ABI.

Oracles: joins/CPU_JOIN/hash_cpu (hash join) and joins/CPU_JOIN/block_cpu (NLJ)
restated in oracle/cpu_join.c, and the sort + equal_range reduction of
tests/common.py.  R tiles and S chunks are set far below the inputs' sizes so
that every test crosses several tiles and chunks, as a relation larger than
device memory would.
"""
import numpy as np
import pytest

import oracle
from eating_one_amd import _abi, gen, EOError, HostArray
from common import brute_join, same_multiset

pytestmark = pytest.mark.gpu


def _outer_want(R, S, inner):
    miss = S[~np.isin(S[:, 0], R[:, 0])]
    nul = np.stack([np.full(len(miss), _abi.EO_NULL_PAYLOAD, np.int32), miss[:, 1]], 1)
    return np.concatenate([inner, nul]).astype(np.int32)


@pytest.mark.parametrize("op", ["hash", "nlj"])
@pytest.mark.parametrize("r_tile,s_chunk", [(0, 0), (70001, 50000), (1 << 15, 1 << 16)])
def test_stream_join_tiles_and_chunks(ctx, op, r_tile, s_chunk):
    nR, nS, kr = (200000, 300000, 1 << 17) if op == "hash" else (90000, 120000, 20000)
    R, S = gen.uniform(11, nR, nS, kr, lo=-(kr // 3))
    want = oracle.hash_join(R, S, nparts=1 << 12)
    st = _abi.StreamStats()
    got = ctx.stream_join(R, S, op=op, r_tile=r_tile, s_chunk=s_chunk, stats=st)
    assert same_multiset(got, want)
    assert st.n_out == len(want)
    rt = r_tile or nR
    sc = s_chunk or nS
    assert st.r_tiles == -(-nR // rt)
    assert st.s_chunks == st.r_tiles * -(-nS // sc)
    assert st.h2d_bytes == 8 * (nR + st.r_tiles * nS)
    assert st.d2h_bytes == 8 * len(want)
    assert ctx.stream_join(R, S, op=op, r_tile=r_tile, s_chunk=s_chunk, count_only=True) == len(want)


def test_stream_join_equals_in_memory_join(ctx):
    """The same inputs through eo_hash_join in HBM and streamed in 5 x 7 pieces."""
    R, S = gen.c3(seed=5, nR=1 << 18, nS=1 << 18)
    whole = ctx.hash_join(R, S)
    assert same_multiset(ctx.stream_join(R, S, r_tile=(1 << 18) // 5 + 1, s_chunk=(1 << 18) // 7 + 1), whole)
    assert same_multiset(ctx.stream_join(R, S, op="nlj", r_tile=(1 << 18) // 5 + 1, s_chunk=(1 << 18) // 7 + 1),
                         whole)


def test_stream_join_pinned_buffers(ctx):
    """eo_host_alloc memory for R, S and out: copied by DMA directly (stats.pinned = 7)."""
    R, S = gen.uniform(12, 150000, 260000, 1 << 16)
    want = brute_join(R, S)
    hR, hS, hO = HostArray(ctx, R.nbytes), HostArray(ctx, S.nbytes), HostArray(ctx, 8 * (len(want) + 16))
    try:
        vR, vS, vO = hR.view(np.int32, R.shape), hS.view(np.int32, S.shape), hO.view(np.int32, (len(want) + 16, 2))
        vR[:] = R
        vS[:] = S
        for op in ("hash", "nlj"):
            st = _abi.StreamStats()
            n, status = ctx.stream_join(vR, vS, op=op, r_tile=40000, s_chunk=70000, out=vO, stats=st)
            assert status == _abi.EO_OK and n == len(want) and st.pinned == 7
            assert same_multiset(vO[:n].copy(), want)
        # pinned input, pageable output and the reverse
        out = np.empty((len(want), 2), np.int32)
        st = _abi.StreamStats()
        n, status = ctx.stream_join(vR, S, r_tile=40000, s_chunk=70000, out=out, stats=st)
        assert status == _abi.EO_OK and st.pinned == 1 and same_multiset(out[:n], want)
        n, status = ctx.stream_join(R, vS, s_chunk=70000, out=vO, stats=st)
        assert status == _abi.EO_OK and st.pinned == 6 and same_multiset(vO[:n].copy(), want)
    finally:
        hO.close()
        hS.close()
        hR.close()


def test_stream_join_output_staging_grows(ctx):
    """Duplicate keys: a chunk yields far more rows than the staging holds, so the
    chunk is joined again into a larger staging (stats.regrown > 0)."""
    R, S = gen.uniform(13, 60000, 90000, 3000)
    want = brute_join(R, S)
    for op in ("hash", "nlj"):
        st = _abi.StreamStats()
        got = ctx.stream_join(R, S, op=op, r_tile=25000, s_chunk=20000, out_chunk=1000, stats=st)
        assert same_multiset(got, want) and st.regrown > 0


def test_stream_join_capacity(ctx):
    """out_cap below the count: EO_ECAPACITY, the full count, and out_cap rows of the result."""
    R, S = gen.uniform(14, 50000, 80000, 20000)
    want = brute_join(R, S)
    cap = len(want) - 1000
    out = np.full((cap, 2), -7, np.int32)
    n, status = ctx.stream_join(R, S, r_tile=20000, s_chunk=30000, out=out)
    assert status == _abi.EO_ECAPACITY and n == len(want)
    # every written row is a row of the join, each used at most as often as it occurs
    rows, cnt = np.unique(out, axis=0, return_counts=True)
    wrows, wcnt = np.unique(want, axis=0, return_counts=True)
    idx = {tuple(r): c for r, c in zip(wrows.tolist(), wcnt.tolist())}
    assert all(idx.get(tuple(r), 0) >= c for r, c in zip(rows.tolist(), cnt.tolist()))


@pytest.mark.parametrize("op", ["hash", "nlj"])
def test_stream_join_result16_and_left_outer(ctx, op):
    R, S = gen.uniform(15, 40000, 70000, 60000)
    inner = brute_join(R, S)
    rows = ctx.stream_join(R, S, op=op, s_chunk=25000, result16=True)
    assert same_multiset(rows[:, [0, 2]], inner)
    assert np.array_equal(rows[:, 1], rows[:, 3])
    got = ctx.stream_join(R, S, op=op, s_chunk=25000, left_outer=True)
    assert same_multiset(got, _outer_want(R, S, inner))
    with pytest.raises(EOError):
        ctx.stream_join(R, S, op=op, r_tile=10000, s_chunk=25000, left_outer=True)  # R in several tiles


def test_stream_join_edge_cases(ctx):
    e = np.empty((0, 2), np.int32)
    R, S = gen.uniform(16, 1000, 3000, 500)
    for op in ("hash", "nlj"):
        assert len(ctx.stream_join(e, S, op=op)) == 0
        assert len(ctx.stream_join(R, e, op=op)) == 0
        got = ctx.stream_join(e, S, op=op, left_outer=True, s_chunk=700)
        assert same_multiset(got, _outer_want(e, S, e))
        r, s = R[:20], S[:30]
        one = ctx.stream_join(r, s, op=op, r_tile=1, s_chunk=1)  # one tuple per tile and chunk
        assert same_multiset(one, oracle.nlj_join(r, s))
        odd = ctx.stream_join(R[:300], S[:500], op=op, r_tile=7, s_chunk=33)
        assert same_multiset(odd, oracle.nlj_join(R[:300], S[:500]))


def test_host_mode_hash_join_streams_large_probe_sides(ctx):
    """eo_hash_join on host buffers with |S| >= 2^22 runs as a streaming join; the
    rows equal the device-resident join's (and the whole-copy path's, taken when
    stats are requested)."""
    R, S = gen.c2(seed=21, nR=1 << 21, nS=(1 << 22) + 3, theta=0.5, selectivity=0.5)
    from eating_one_amd import DeviceArray
    dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
    n, _ = ctx.hash_join_device(dR, len(R), dS, len(S), None, 0)
    dout = DeviceArray(ctx, 8 * n)
    n2, st = ctx.hash_join_device(dR, len(R), dS, len(S), dout, n)
    assert st == _abi.EO_OK and n2 == n
    want = dout.to_host(np.int32, (n, 2))
    assert same_multiset(ctx.hash_join(R, S), want)
    assert same_multiset(ctx.hash_join(R, S, stats=_abi.JoinStats()), want)
    rows = ctx.hash_join(R, S, result16=True)
    assert same_multiset(rows[:, [0, 2]], want)
    outer = ctx.hash_join(R, S, left_outer=True)
    assert same_multiset(outer, _outer_want(R, S, want))
    for a in (dout, dS, dR):
        a.close()
