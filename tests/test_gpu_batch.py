"""GPU parity of resident tables and batched joins (persistent probe kernel).

Every request of a batch must give the oracle's result (same count, same
sorted multiset of pairs), which is also what the request gives alone through
eo_hash_join; graph replays must repeat the eager result every time.
"""
import ctypes

import numpy as np
import pytest

import eating_one_amd as eo
import oracle
from eating_one_amd import DeviceArray, gen
from common import brute_join, same_multiset

pytestmark = pytest.mark.gpu


def _oracle(R, S):
    return oracle.hash_join(R, S, nparts=max(1, len(R) // 64))


def _mixed_inputs(seed, n_req, lo=10, hi=20):
    """Sizes log-uniform in [2^lo, 2^hi]; keys uniform; selectivity U[0.1, 1]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_req):
        nR = int(2 ** rng.uniform(lo, hi))
        nS = int(2 ** rng.uniform(lo, hi))
        sel = float(rng.uniform(0.1, 1.0))
        R, S = gen.c2(seed=seed * 1000 + i, nR=nR, nS=nS, selectivity=sel)
        out.append((R, S))
    return out


class _Req:
    """Device buffers of one request."""

    def __init__(self, ctx, R, S, cap=None, table=None, s_offset=0, want=None):
        self.R, self.S = R, S
        self.want = _oracle(R, S) if want is None else want
        self.cap = len(self.want) if cap is None else cap
        self.dR = None if table is not None else DeviceArray.from_host(ctx, R)
        # s_offset tuples of padding in front of S: a device pointer that is only 8-byte aligned
        pad = np.zeros((s_offset, 2), dtype=np.int32)
        self.dS = DeviceArray.from_host(ctx, np.concatenate([pad, S]) if s_offset else S)
        self.s_ptr = self.dS.ptr + 8 * s_offset
        self.dout = DeviceArray(ctx, max(self.cap, 1) * 8) if self.cap else None
        self.table = table

    def request(self, tag, count_only=False):
        return eo.request(R=self.dR, nR=len(self.R), S=self.s_ptr, nS=len(self.S),
                          out=None if count_only or not self.dout else self.dout, out_cap=0 if count_only else self.cap,
                          table=self.table, tag=tag)

    def pairs(self, n):
        return self.dout.to_host(np.int32, (n, 2)) if n else np.empty((0, 2), np.int32)


def test_batch_mixed_sizes_equal_oracle_and_solo(ctx):
    inputs = _mixed_inputs(51, 24)
    reqs = [_Req(ctx, R, S) for R, S in inputs]
    b = ctx.batch()
    res = b.run([r.request(tag=100 + i) for i, r in enumerate(reqs)])
    for i, (r, (tag, n, st)) in enumerate(zip(reqs, res)):
        assert tag == 100 + i
        assert st == 0
        assert n == len(r.want), (i, len(r.R), len(r.S))
        got = r.pairs(n)
        assert same_multiset(got, r.want), i
        # the same request alone
        solo = ctx.hash_join(r.R, r.S)
        assert same_multiset(solo, got), i
    b.free()


def test_batch_hot_partition_tiles_beyond_descriptors(ctx):
    """Keys that are all multiples of 2^12 land in one partition: that join has
    far more probe tiles (R chunks x S spans) than the evenly-spread estimate
    the batch writes descriptors for, so its later tiles resolve by search in
    the probe; every result still equals the oracle's and the solo join's."""
    rng = np.random.default_rng(77)
    nR, nS = 1 << 17, 1 << 20  # 8 R chunks x 8 S spans = 64 tiles; 16 estimated
    rk = (rng.permutation(nR).astype(np.int64) << 12).astype(np.int32)
    R = np.stack([rk, rng.integers(0, 1 << 30, nR).astype(np.int32)], 1)
    S = np.stack([rk[rng.integers(0, nR, nS)], np.arange(nS, dtype=np.int32)], 1)
    # the oracle's nested loop over one hot partition is quadratic: sort + equal_range instead
    reqs = [_Req(ctx, R, S, want=brute_join(R, S))]
    reqs += [_Req(ctx, r, s) for r, s in _mixed_inputs(78, 6, lo=12, hi=16)]  # their spare estimate: < 30 tiles
    b = ctx.batch()
    res = b.run([r.request(tag=i) for i, r in enumerate(reqs)])
    for i, (r, (tag, n, st)) in enumerate(zip(reqs, res)):
        assert st == 0 and n == len(r.want), i
        got = r.pairs(n)
        assert same_multiset(got, r.want), i
        if i == 0:
            assert n == nS
            assert same_multiset(ctx.hash_join(r.R, r.S), got)
    b.free()


def test_batch_small_probe_sides_build_on_s(ctx):
    """Joins whose S is smaller than R build their LDS tables on S in a batch
    (and the partition bits follow S); rows come out as (R payload, S
    payload) and equal the oracle's and the solo join's, beside ordinary
    joins in the same batch."""
    inputs = [gen.c2(seed=301, nR=(1 << 22) + 77, nS=1 << 18, selectivity=0.7),
              gen.c2(seed=302, nR=3 << 21, nS=(1 << 20) + 5, selectivity=1.0),
              gen.c2(seed=303, nR=1 << 21, nS=1 << 21, selectivity=0.5)]
    inputs += _mixed_inputs(304, 4, lo=12, hi=18)
    reqs = [_Req(ctx, R, S) for R, S in inputs]
    b = ctx.batch()
    res = b.run([r.request(tag=i) for i, r in enumerate(reqs)])
    for i, (r, (tag, n, st)) in enumerate(zip(reqs, res)):
        assert st == 0 and n == len(r.want), i
        got = r.pairs(n)
        assert same_multiset(got, r.want), i
        if i < 3:
            assert same_multiset(ctx.hash_join(r.R, r.S), got), i
    b.free()


def test_batch_resident_tables_and_count_only(ctx):
    rng = np.random.default_rng(7)
    tabs = []
    for k, nR in enumerate([3000, 40000, 300000]):
        R, _ = gen.c2(seed=70 + k, nR=nR, nS=1)
        tabs.append((R, ctx.table_create(R)))
    assert tabs[0][1].info()["radix_bits"] == 0
    assert tabs[2][1].info()["radix_bits"] > 0
    reqs = []
    for i in range(9):
        R, t = tabs[i % 3]
        nS = int(2 ** rng.uniform(8, 19))
        keys = rng.integers(0, 2 * len(R), size=nS).astype(np.int32)
        S = np.stack([keys, np.arange(nS, dtype=np.int32)], 1)
        reqs.append(_Req(ctx, R, S, table=t))
    b = ctx.batch()
    res = b.run([r.request(tag=i, count_only=(i % 4 == 3)) for i, r in enumerate(reqs)])
    for i, (r, (tag, n, st)) in enumerate(zip(reqs, res)):
        assert st == 0 and n == len(r.want), i
        if i % 4 != 3:
            assert same_multiset(r.pairs(n), r.want), i
    for _, t in tabs:
        t.free()
    b.free()


def test_batch_edge_cases(ctx):
    """Empty sides, capacity overflow, a misaligned S (individual stable
    partition), a forced two-pass request and duplicate-heavy keys."""
    R1, S1 = gen.c2(seed=901, nR=50000, nS=70000)
    R2, S2 = gen.uniform(902, 30000, 40000, key_range=1 << 10)  # ~1.2M pairs, partitions overflow one LDS table
    R3, S3 = gen.c2(seed=903, nR=100000, nS=5000)
    empty = np.empty((0, 2), np.int32)
    reqs = [
        _Req(ctx, R1, S1),
        _Req(ctx, R2, S2),
        _Req(ctx, R1, S1, cap=1000),           # too small: EO_ECAPACITY, full count
        _Req(ctx, R3, S3, s_offset=1),         # S only 8-byte aligned
        _Req(ctx, R1[:1], S1[:10]),
    ]
    b = ctx.batch()
    q = [r.request(tag=i) for i, r in enumerate(reqs)]
    # empty relations
    q.append(eo.request(R=reqs[0].dR, nR=0, S=reqs[0].dS, nS=len(S1), tag=99))
    q.append(eo.request(R=reqs[0].dR, nR=len(R1), S=reqs[0].dS, nS=0, tag=98))
    res = b.run(q)
    assert res[5] == (99, 0, 0) and res[6] == (98, 0, 0)
    for i in (0, 1, 3, 4):
        tag, n, st = res[i]
        assert st == 0 and n == len(reqs[i].want), i
        assert same_multiset(reqs[i].pairs(n), reqs[i].want), i
    tag, n, st = res[2]
    assert st == eo._abi.EO_ECAPACITY and n == len(reqs[2].want)
    # forced two radix passes (14 bits) for every request without a table
    res2 = b.run(q, radix_bits=14)
    for i in (0, 1, 3):
        tag, n, st = res2[i]
        assert st == 0 and n == len(reqs[i].want), i
        assert same_multiset(reqs[i].pairs(n), reqs[i].want), i
    b.free()


def test_batch_rejects_bad_requests(ctx):
    b = ctx.batch()
    bad = eo.request(R=None, nR=10, S=None, nS=10)
    with pytest.raises(eo.EOError):
        b.run([bad])
    # the batch object is still usable
    R, S = gen.c2(seed=5, nR=2000, nS=3000)
    r = _Req(ctx, R, S)
    (tag, n, st), = b.run([r.request(tag=1)])
    assert n == len(r.want)
    b.free()


def test_batch_replay_refused_after_failed_submit(ctx):
    """A replay after a submit that failed (here: a request above the row limit) launches nothing."""
    R, S = gen.c2(seed=6, nR=2000, nS=3000)
    r = _Req(ctx, R, S)
    too_big = eo.request(R=r.dR, nR=len(R), S=r.s_ptr, nS=1 << 31, out=r.dout, out_cap=r.cap)
    # a fresh batch object whose first submit fails
    b = ctx.batch()
    with pytest.raises(eo.EOError):
        b.submit([too_big])
    assert ctx.lib.eo_batch_replay(b.handle) == eo._abi.EO_EINVAL
    # a good batch, then a failed submit on the same object: the old graph is gone too
    assert b.run([r.request(tag=1)])[0][1] == len(r.want)
    b.replay()
    assert b.wait()[0][1] == len(r.want)
    with pytest.raises(eo.EOError):
        b.submit([r.request(tag=2), too_big])
    assert ctx.lib.eo_batch_replay(b.handle) == eo._abi.EO_EINVAL
    # and the object still works
    assert b.run([r.request(tag=3)])[0][1] == len(r.want)
    b.free()


def test_batch_graph_replay_matches_eager(ctx):
    inputs = _mixed_inputs(61, 10, lo=10, hi=18)
    reqs = [_Req(ctx, R, S) for R, S in inputs]
    b = ctx.batch()
    eager = b.run([r.request(tag=i) for i, r in enumerate(reqs)])
    eager_pairs = [r.pairs(n) for r, (_, n, _) in zip(reqs, eager)]
    R0, S0 = gen.c2(seed=62, nR=5000, nS=9000)
    for rep in range(4):
        b.replay()
        res = b.wait()
        assert res == eager, rep
        for r, (_, n, _), want in zip(reqs, res, eager_pairs):
            assert same_multiset(r.pairs(n), want), rep
        # eager work on the same device between replays
        assert len(ctx.hash_join(R0, S0)) == len(_oracle(R0, S0))
    # new S contents in the same buffers: the replay sees them
    rng = np.random.default_rng(63)
    for r in reqs:
        keys = rng.integers(0, len(r.R), size=len(r.S)).astype(np.int32)
        S2 = np.stack([keys, np.arange(len(r.S), dtype=np.int32)], 1)
        r.S = S2
        r.want = _oracle(r.R, S2)
        ctx.lib.eo_copy_h2d(ctx.handle, r.dS.ptr, ctypes.c_void_p(S2.ctypes.data), S2.nbytes)
    b.replay()
    res = b.wait()
    for i, (r, (_, n, st)) in enumerate(zip(reqs, res)):
        if len(r.want) > r.cap:
            assert st == eo._abi.EO_ECAPACITY and n == len(r.want)
        else:
            assert st == 0 and n == len(r.want)
            assert same_multiset(r.pairs(n), r.want), i
    b.free()


def test_batch_stats_and_footprint(ctx):
    inputs = _mixed_inputs(71, 6, lo=12, hi=18)
    reqs = [_Req(ctx, R, S) for R, S in inputs]
    q = [r.request(tag=i) for i, r in enumerate(reqs)]
    b = ctx.batch()
    fp = b.footprint(q)
    # a join whose smaller side passes one LDS table partitions both relations
    assert fp > sum(8 * (len(r.R) + len(r.S)) for r in reqs if min(len(r.R), len(r.S)) > 16384)
    st = eo.JoinStats()
    res = b.run(q, stats=st)
    assert st.n_out == sum(n for _, n, _ in res)
    names = st.kernel_table()
    assert "batch_probe" in names
    assert st.ms_total > 0
    b.free()


def test_batch_result16_and_left_outer(ctx):
    """Batch flags: EO_F_RESULT16 rows and EO_F_LEFT_OUTER null rows for every
    request (resident table, empty build side, misaligned S included); graph
    replays repeat the eager rows."""
    from eating_one_amd import _abi
    inputs = _mixed_inputs(71, 6, lo=10, hi=17)
    R0, S0 = gen.c2(seed=72, nR=5000, nS=9000, selectivity=0.4)
    table = ctx.table_create(R0)
    empty = np.empty((0, 2), np.int32)
    cases = [(R, S, None, 0) for R, S in inputs] + [(R0, S0, table, 0), (empty, S0, None, 0), (R0, S0, None, 1)]

    def outer_want(R, S):
        inner = _oracle(R, S) if len(R) else np.empty((0, 2), np.int32)
        miss = S[~np.isin(S[:, 0], R[:, 0])]
        return np.concatenate([inner, np.stack([np.full(len(miss), _abi.EO_NULL_PAYLOAD, np.int32), miss[:, 1]], 1)])

    b = ctx.batch()
    try:
        for wide in (False, True):
            reqs, bufs = [], []
            for i, (R, S, t, off) in enumerate(cases):
                want = outer_want(R, S)
                row = 16 if wide else 8
                dR = DeviceArray.from_host(ctx, R) if (t is None and len(R)) else None
                pad = np.zeros((off, 2), np.int32)
                dS = DeviceArray.from_host(ctx, np.concatenate([pad, S]) if off else S)
                dout = DeviceArray(ctx, max(len(want), 1) * row)
                bufs.append((dR, dS, dout, want))
                reqs.append(eo.request(R=dR, nR=len(R), S=dS.ptr + 8 * off, nS=len(S), out=dout, out_cap=len(want),
                                       table=t, tag=i))
            for rep in range(2):
                res = b.run(reqs, left_outer=True, result16=wide) if rep == 0 else (b.replay() or b.wait())
                for (tag, n, st), (dR, dS, dout, want) in zip(res, bufs):
                    assert st == _abi.EO_OK and n == len(want), (tag, n, len(want))
                    rows = dout.to_host(np.int32, (n, 4 if wide else 2))
                    got = rows[:, [0, 2]] if wide else rows
                    assert same_multiset(got, want), tag
            for dR, dS, dout, _ in bufs:
                for a in (dout, dS, dR):
                    if a is not None:
                        a.close()
    finally:
        b.free()
        table.free()


def test_batch_sampled_partitions_equal_oracle_and_replay(ctx):
    """Relations of 2^20 tuples or more with up to 8 radix bits, in an inner
    batch of 256 requests or more, take one-pass partitions sized from a 1/16
    sample (runs reserved with atomics, gaps between partitions): every
    request equals the oracle and its solo join, beside small requests, and
    graph replays repeat the rows."""
    inputs = [gen.c2(seed=401, nR=(1 << 20) + 333, nS=(1 << 21) + 17, selectivity=0.8),
              gen.c2(seed=402, nR=(1 << 21) + 5, nS=(1 << 20) + 9, selectivity=1.0),  # built on S
              gen.c2(seed=403, nR=1 << 20, nS=1 << 22, theta=0.9)]
    inputs += _mixed_inputs(404, 6, lo=10, hi=18) + _mixed_inputs(406, 250, lo=6, hi=10)
    reqs = [_Req(ctx, R, S) for R, S in inputs]
    st = eo.JoinStats()
    b = ctx.batch()
    res = b.run([r.request(tag=i) for i, r in enumerate(reqs)], stats=st)
    assert st.sampled == 1
    eager = []
    for i, (r, (tag, n, s)) in enumerate(zip(reqs, res)):
        assert s == 0 and n == len(r.want), (i, len(r.R), len(r.S))
        got = r.pairs(n)
        assert same_multiset(got, r.want), i
        if i < 3:
            assert same_multiset(ctx.hash_join(r.R, r.S), got), i
        eager.append(got)
    for _ in range(2):
        b.replay()
        res2 = b.wait()
        for i, (r, (tag, n, s)) in enumerate(zip(reqs, res2)):
            assert s == 0 and n == len(r.want), i
            assert same_multiset(r.pairs(n), eager[i]), i
    b.free()


def _sampled_line(t, j):
    """Line of tile t that the partition sample reads for group j (restated to
    build an input that defeats the sample)."""
    h = ((t * 64 + j) * 0x9E3779B1) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x85EBCA77) & 0xFFFFFFFF
    h ^= h >> 13
    return j * 16 + (h >> 28)


def test_batch_sampled_partition_overflow_redoes_the_batch_exactly(ctx):
    """One request's S holds keys spread over all partitions in every line the
    sample reads and one hot key everywhere else: its hot partition's region
    overflows, and waiting for the batch redoes it with exact partitions
    (stats.sampled == 2); every request still equals the oracle.  A left
    outer batch keeps exact partitions."""
    rng = np.random.default_rng(0x5E)
    n = 1 << 21
    keys = np.full(n, 5, dtype=np.int32)
    for t in range(n // 16384):
        for j in range(64):
            base = t * 16384 + _sampled_line(t, j) * 16
            keys[base:base + 16] = rng.integers(0, 1 << 20, size=16, dtype=np.int32)
    S = np.stack([keys, np.arange(n, dtype=np.int32)], 1)
    R = np.stack([rng.permutation(1 << 20).astype(np.int32), np.arange(1 << 20, dtype=np.int32)], 1)
    inputs = [(R, S)] + _mixed_inputs(405, 4, lo=12, hi=16) + _mixed_inputs(407, 251, lo=6, hi=10)
    reqs = [_Req(ctx, r, s) for r, s in inputs]
    st = eo.JoinStats()
    b = ctx.batch()
    res = b.run([r.request(tag=i) for i, r in enumerate(reqs)], stats=st)
    assert st.sampled == 2
    for i, (r, (tag, n_out, s)) in enumerate(zip(reqs, res)):
        assert s == 0 and n_out == len(r.want), i
        assert same_multiset(r.pairs(n_out), r.want), i
    # replays of the batch (captured from its exact redo) repeat every row
    for _ in range(2):
        b.replay()
        res2 = b.wait()
        for i, (r, (tag, n_out, s)) in enumerate(zip(reqs, res2)):
            assert s == 0 and n_out == len(r.want), i
            assert same_multiset(r.pairs(n_out), r.want), i
    # a left outer batch, and a batch of few requests, keep exact partitions
    st2 = eo.JoinStats()
    b.run([r.request(tag=i, count_only=True) for i, r in enumerate(reqs)], stats=st2, left_outer=True)
    assert st2.sampled == 0
    st3 = eo.JoinStats()
    res = b.run([r.request(tag=i) for i, r in enumerate(reqs[:5])], stats=st3)
    assert st3.sampled == 0
    for i, (r, (tag, n_out, s)) in enumerate(zip(reqs[:5], res)):
        assert s == 0 and n_out == len(r.want), i
        assert same_multiset(r.pairs(n_out), r.want), i
    b.free()


@pytest.mark.parametrize("run,sampled", [(2, 1), (8, 2)])
def test_batch_sampled_clumping(ctx, run, sampled):
    """A large relation whose S keys come in runs of equal keys: the batch
    sample measures the clumping.  Pairs (variance ratio ~2) keep sampled
    partitions with widened margins; runs of 8 exceed what the relation's
    buffer is sized for, the layout raises the capacity flag and waiting for
    the batch redoes it exactly.  Every request equals the oracle."""
    rng = np.random.default_rng(0x70 + run)
    nR, nS = 1 << 20, (1 << 21) + 37
    R = np.stack([rng.permutation(nR).astype(np.int32), np.arange(nR, dtype=np.int32)], 1)
    skeys = np.repeat(rng.integers(0, nR, nS // run + 1).astype(np.int32), run)[:nS]
    S = np.stack([skeys, np.arange(nS, dtype=np.int32)], 1)
    inputs = [(R, S)] + _mixed_inputs(0x71 + run, 255, lo=6, hi=10)
    reqs = [_Req(ctx, r, s) for r, s in inputs]
    st = eo.JoinStats()
    b = ctx.batch()
    try:
        res = b.run([r.request(tag=i) for i, r in enumerate(reqs)], stats=st)
        assert st.sampled == sampled
        for i, (r, (tag, n_out, s)) in enumerate(zip(reqs, res)):
            assert s == 0 and n_out == len(r.want), i
            assert same_multiset(r.pairs(n_out), r.want), i
    finally:
        b.free()
