"""GPU parity of the primitives and the hash / nested-loop joins, through the C ABI.

Every result is compared with the CPU oracle on the same seeded inputs:
bit-exact for scans and stable partitions, as a sorted multiset with an
identical count for joins.
"""
import numpy as np
import pytest

import eating_one_amd as eo
import oracle
from eating_one_amd import CapacityError, _abi, gen
from common import (brute_join, eo_to_ref, join_count_closed_form, join_ori_doc, libc_rand_table, ref_to_eo,
                    same_multiset, sha256)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 4095, 4096, 4097, 100003, (1 << 20) - 1, 1 << 20, (1 << 20) + 16383,
                               (1 << 22) + 17])
def test_scan_u32(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1 << 10, size=n).astype(np.uint32)
    out, tot = ctx.exclusive_scan(a)
    want, wtot = oracle.exclusive_scan(a)
    assert np.array_equal(out, want)
    assert tot == wtot


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 300001, (1 << 20) + 8191, 1 << 21])
def test_scan_u64(ctx, n):
    rng = np.random.default_rng(n + 1)
    a = rng.integers(0, 1 << 40, size=n).astype(np.uint64)
    out, tot = ctx.exclusive_scan(a)
    want, wtot = oracle.exclusive_scan(a)
    assert np.array_equal(out, want)
    assert tot == wtot


def test_partition_reproduces_join_ori_hash_part_files(ctx):
    doc = join_ori_doc()
    for table in ("R", "S"):
        T = ref_to_eo(libc_rand_table(doc["tables"][table]["rows"]))
        out, off = ctx.radix_partition(T, 2)
        for p in range(4):
            part = eo_to_ref(out[off[p]:off[p + 1]])
            assert sha256(part) == doc["partitions"][f"{table}-{p}"]["sha256"]


@pytest.mark.parametrize("bits,shift", [(1, 0), (4, 0), (6, 3), (8, 0), (9, 0), (11, 0), (11, 5), (12, 0), (14, 0),
                                        (16, 0), (8, 24)])
def test_partition_stable_bits(ctx, bits, shift):
    rng = np.random.default_rng(bits * 100 + shift)
    n = 200003
    T = np.stack([rng.integers(-(1 << 31), (1 << 31) - 1, size=n, dtype=np.int64).astype(np.int32),
                  np.arange(n, dtype=np.int32)], 1)
    out, off = ctx.radix_partition(T, bits, shift)
    want, woff = oracle.partition_bits(T, bits, shift)
    assert np.array_equal(off, woff)
    assert np.array_equal(out, want)


def test_partition_skewed_and_tiny(ctx):
    for n in (0, 1, 5, 4096, 4097):
        T = np.stack([np.full(n, 7, dtype=np.int32), np.arange(n, dtype=np.int32)], 1)
        out, off = ctx.radix_partition(T, 8)
        want, woff = oracle.partition_bits(T, 8, 0)
        assert np.array_equal(out, want) and np.array_equal(off, woff)


@pytest.mark.parametrize("bits", [0, 12])
def test_hash_join_misaligned_device_input(ctx, bits):
    """Device pointers that are only 8-byte aligned take the stable scatter."""
    R, S = gen.c3(seed=21, nR=70001, nS=50003, key_range=1 << 14)
    import ctypes
    from eating_one_amd import DeviceArray, _abi
    dR = DeviceArray.from_host(ctx, np.concatenate([np.zeros((1, 2), np.int32), R]))
    dS = DeviceArray.from_host(ctx, np.concatenate([np.zeros((1, 2), np.int32), S]))
    want = oracle.hash_join(R, S, nparts=1 << 10)
    dout = DeviceArray(ctx, len(want) * 8 + 8)
    n = ctypes.c_uint64()
    o = ctx.opts(where=_abi.EO_DEVICE, radix_bits=bits)
    st = ctx.lib.eo_hash_join(ctx.handle, dR.ptr + 8, len(R), dS.ptr + 8, len(S), dout.ptr, len(want),
                              ctypes.byref(n), ctypes.byref(o))
    assert st == 0 and n.value == len(want)
    got = dout.to_host(np.int32, (len(want), 2), len(want) * 8)
    assert same_multiset(got, want)


def test_hash_join_c1_full(ctx):
    R, S = gen.c1()
    got = ctx.hash_join(R, S)
    assert len(got) == 1 << 20
    assert same_multiset(got, oracle.hash_join(R, S))


@pytest.mark.parametrize("det", [False, True])
@pytest.mark.parametrize("bits,passes,cap", [(-1, 0, 16384), (3, 1, 4096), (8, 1, 8192), (11, 1, 4096),
                                             (12, 2, 4096), (14, 2, 16384), (0, 0, 8192), (2, 2, 16384),
                                             (17, 2, 16384), (22, 2, 16384)])
def test_hash_join_plans(ctx, bits, passes, cap, det):
    R, S = gen.c3(seed=bits + 100, nR=60000, nS=90000, key_range=1 << 15)
    got = ctx.hash_join(R, S, radix_bits=bits, passes=passes, lds_tuples=cap, deterministic=det)
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert len(got) == len(want) == join_count_closed_form(R, S)
    assert same_multiset(got, want)


@pytest.mark.parametrize("bits,passes,cap", [(-1, 0, 16384), (5, 1, 16384), (8, 1, 8192), (11, 1, 4096),
                                             (12, 2, 4096), (13, 2, 16384)])
def test_hash_join_deterministic_order_repeats(ctx, bits, passes, cap):
    """EO_F_DETERMINISTIC: the same rows in the same order on every run, with
    duplicate keys on both sides (partitions with positions fixed by the
    input — ordered fast passes up to 7 bits per pass, else stable ones — and
    a fixed bucket order)."""
    R, S = gen.c3(seed=bits + 7, nR=70000, nS=110000, key_range=1 << 14)
    R = np.concatenate([R, R[:5000]])  # more duplicates, some several LDS chunks deep
    runs = [ctx.hash_join(R, S, radix_bits=bits, passes=passes, lds_tuples=cap, deterministic=True)
            for _ in range(3)]
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])
    assert same_multiset(runs[0], oracle.hash_join(R, S, nparts=1 << 12))


def test_hash_join_edge_cases(ctx):
    e = np.empty((0, 2), dtype=np.int32)
    one = np.array([[5, 9]], dtype=np.int32)
    assert len(ctx.hash_join(e, e)) == 0
    assert len(ctx.hash_join(one, e)) == 0
    assert len(ctx.hash_join(e, one)) == 0
    assert same_multiset(ctx.hash_join(one, one), np.array([[9, 9]]))
    # extreme and negative keys
    R = np.array([[-(1 << 31), 0], [(1 << 31) - 1, 1], [0, 2], [-1, 3]], dtype=np.int32)
    S = np.array([[(1 << 31) - 1, 0], [-(1 << 31), 1], [-1, 2], [7, 3]], dtype=np.int32)
    assert same_multiset(ctx.hash_join(R, S), np.array([[1, 0], [0, 1], [3, 2]]))
    assert same_multiset(ctx.hash_join(R, S, radix_bits=4), np.array([[1, 0], [0, 1], [3, 2]]))


@pytest.mark.parametrize("det", [False, True])
def test_hash_join_heavy_duplicates(ctx, det):
    # one key everywhere: R partition larger than any LDS table -> several R chunks
    R = np.stack([np.full(20000, 42), np.arange(20000)], 1).astype(np.int32)
    S = np.stack([np.full(300, 42), np.arange(300)], 1).astype(np.int32)
    got = ctx.hash_join(R, S, radix_bits=4, deterministic=det)
    assert len(got) == 20000 * 300
    assert same_multiset(got, brute_join(R, S))


def test_hash_join_overfull_subpartition_rounds(ctx):
    """One key fills a whole partition: its R side spans many LDS tables."""
    R = np.stack([np.full(1 << 17, 5), np.arange(1 << 17)], 1).astype(np.int32)
    extra = np.stack([np.arange(1 << 15) * 4 + 2, (1 << 17) + np.arange(1 << 15)], 1).astype(np.int32)
    R = np.concatenate([R, extra])
    S = np.stack([np.concatenate([np.full(64, 5), np.arange(1000) * 4 + 2]), np.arange(1064)], 1).astype(np.int32)
    got = ctx.hash_join(R, S, radix_bits=2)
    assert len(got) == 64 * (1 << 17) + 1000
    assert same_multiset(got, brute_join(R, S))


@pytest.mark.parametrize("det", [False, True])
def test_hash_join_count_only_and_capacity(ctx, det):
    R, S = gen.c3(seed=5, nR=5000, nS=5000, key_range=1000)
    n = ctx.hash_join(R, S, count_only=True, deterministic=det)
    assert n == join_count_closed_form(R, S)
    import ctypes
    from eating_one_amd import _abi
    out = np.empty((n - 1, 2), dtype=np.int32)
    cnt = ctypes.c_uint64()
    o = ctx.opts(deterministic=det)
    st = ctx.lib.eo_hash_join(ctx.handle, R.ctypes.data, len(R), S.ctypes.data, len(S), out.ctypes.data, n - 1,
                              ctypes.byref(cnt), ctypes.byref(o))
    assert st == _abi.EO_ECAPACITY and cnt.value == n


@pytest.mark.parametrize("theta,sel", [(0.0, 1.0), (0.5, 1.0), (0.9, 1.0), (0.0, 0.1), (0.9, 0.1)])
def test_hash_join_c2_shape_small(ctx, theta, sel):
    R, S = gen.c2(seed=17, nR=1 << 18, nS=1 << 19, theta=theta, selectivity=sel)
    got = ctx.hash_join(R, S)
    want = oracle.hash_join(R, S, nparts=1 << 12)
    assert same_multiset(got, want)


@pytest.mark.parametrize("bits,theta,sel", [(13, 0.0, 1.0), (13, 0.9, 1.0), (14, 0.5, 0.1), (9, 0.0, 1.0), (5, 0.9, 1.0)])
def test_hash_join_two_pass_shapes(ctx, bits, theta, sel):
    """Two MSD passes from one combined-digit histogram, at C2's shape scaled down."""
    R, S = gen.c2(seed=bits * 7 + 3, nR=1 << 21, nS=1 << 22, theta=theta, selectivity=sel)
    st = eo.JoinStats()
    got = ctx.hash_join(R, S, radix_bits=bits, passes=2, stats=st)
    assert st.passes == 2 and st.radix_bits == bits
    assert "seg_hist" not in st.kernel_table()
    want = oracle.hash_join(R, S, nparts=1 << 15)
    assert len(got) == len(want)
    assert same_multiset(got, want)


def test_nlj_small_and_edges(ctx):
    R, S = gen.c3(seed=9, nR=3000, nS=5000, key_range=200)
    assert same_multiset(ctx.nlj_join(R, S), oracle.nlj_join(R, S))
    e = np.empty((0, 2), dtype=np.int32)
    assert len(ctx.nlj_join(e, R)) == 0
    # ragged tiles: sizes not multiples of the 4096 x 2048 tile
    R, S = gen.c3(seed=10, nR=4097, nS=2049, key_range=50)
    assert same_multiset(ctx.nlj_join(R, S), oracle.nlj_join(R, S))


def test_nlj_fused_modes_and_dense_hits(ctx):
    """Fused pass vs deterministic count/scan/write; every R group hitting; capacity."""
    R, S = gen.c3(seed=12, nR=9000, nS=7000, key_range=300)
    want = oracle.nlj_join(R, S)
    fused = ctx.nlj_join(R, S)
    det1 = ctx.nlj_join(R, S, deterministic=True)
    det2 = ctx.nlj_join(R, S, deterministic=True)
    assert same_multiset(fused, want) and same_multiset(det1, want)
    assert np.array_equal(det1, det2)
    assert ctx.nlj_join(R, S, count_only=True) == len(want)
    # one key everywhere: every group of every wave records a hit
    R1 = np.stack([np.full(4100, 7), np.arange(4100)], 1).astype(np.int32)
    S1 = np.stack([np.full(700, 7), np.arange(700)], 1).astype(np.int32)
    got = ctx.nlj_join(R1, S1)
    assert len(got) == 4100 * 700
    assert same_multiset(got, brute_join(R1, S1))
    import ctypes
    from eating_one_amd import _abi
    out = np.empty((len(want) - 1, 2), dtype=np.int32)
    cnt = ctypes.c_uint64()
    o = ctx.opts()
    st = ctx.lib.eo_nlj_join(ctx.handle, R.ctypes.data, len(R), S.ctypes.data, len(S), out.ctypes.data,
                             len(want) - 1, ctypes.byref(cnt), ctypes.byref(o))
    assert st == _abi.EO_ECAPACITY and cnt.value == len(want)


def test_nlj_c3_full(ctx):
    R, S = gen.c3()
    got = ctx.nlj_join(R, S)
    assert len(got) == join_count_closed_form(R, S)
    assert same_multiset(got, oracle.hash_join(R, S, nparts=1 << 12))


@pytest.mark.parametrize("det", [False, True])
def test_result16_rows_all_joins(ctx, det):
    """EO_F_RESULT16: rows {R payload, key, S payload, key} as the reference's RESULT."""
    R, S = gen.c3(seed=91, nR=40000, nS=60000, key_range=15000)
    want = oracle.hash_join(R, S, nparts=1 << 10)
    rkey = dict(zip(R[:, 1].tolist(), R[:, 0].tolist()))  # payload -> key
    for name, fn in (("hash", lambda **k: ctx.hash_join(R, S, **k)), ("nlj", lambda **k: ctx.nlj_join(R, S, **k))):
        rows = fn(result16=True, deterministic=det)
        assert rows.shape == (len(want), 4), name
        assert np.array_equal(rows[:, 1], rows[:, 3]), name
        assert same_multiset(rows[:, [0, 2]], want), name
        assert all(rkey[int(p)] == int(k) for p, k in zip(rows[:2000, 0], rows[:2000, 1])), name
    ix = ctx.sidx_build(R, learned=True)
    rows = ix.join(S, result16=True, deterministic=det)
    assert rows.shape == (len(want), 4) and np.array_equal(rows[:, 1], rows[:, 3])
    assert same_multiset(rows[:, [0, 2]], want)
    ix.free()


def test_zero_copy_host_output(ctx):
    """Rows written by the kernels straight into pinned, device-mapped host memory."""
    from eating_one_amd import DeviceArray, HostArray
    R, S = gen.c2(seed=95, nR=1 << 16, nS=1 << 18, selectivity=0.5)
    want = oracle.hash_join(R, S, nparts=1 << 10)
    dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
    h = HostArray(ctx, len(want) * 8)
    n, st = ctx.hash_join_device(dR, len(R), dS, len(S), h, len(want))
    assert st == 0 and n == len(want)
    assert same_multiset(h.view(np.int32, (n, 2)).copy(), want)
    h.close()


def _outer_want(R, S, inner):
    """Left outer join = the oracle's inner join plus (NULL, S payload) for every
    S tuple whose key is absent from R (set membership, numpy.isin)."""
    from eating_one_amd import _abi
    miss = S[~np.isin(S[:, 0], R[:, 0])]
    nul = np.stack([np.full(len(miss), _abi.EO_NULL_PAYLOAD, np.int32), miss[:, 1]], 1)
    return np.concatenate([inner, nul]).astype(np.int32)


@pytest.mark.parametrize("bits", [-1, 4, 13])
def test_left_outer_hash_join(ctx, bits):
    R, S = gen.c2(seed=97 + max(bits, 0), nR=1 << 17, nS=1 << 18, theta=0.5, selectivity=0.4)
    want = _outer_want(R, S, oracle.hash_join(R, S, nparts=1 << 10))
    kw = dict(radix_bits=bits, passes=2 if bits > 11 else 0)
    got = ctx.hash_join(R, S, left_outer=True, **kw)
    assert same_multiset(got, want)
    assert ctx.hash_join(R, S, left_outer=True, count_only=True, **kw) == len(want)
    rows = ctx.hash_join(R, S, left_outer=True, result16=True, **kw)
    assert same_multiset(rows[:, [0, 2]], want)
    nul = rows[:, 0] == -(1 << 31)
    assert (rows[nul, 1] == -(1 << 31)).all() and np.array_equal(rows[:, 2] >= 0, np.ones(len(rows), bool))


def test_left_outer_multi_chunk_partitions_and_empty_build(ctx):
    # one key fills partitions beyond one LDS table: S tuples probe several R chunks
    R = np.stack([np.full(40000, 3), np.arange(40000)], 1).astype(np.int32)
    S = np.stack([np.array([3, 4, 5, 3] * 50), np.arange(200)], 1).astype(np.int32)
    want = _outer_want(R, S, brute_join(R, S))
    assert same_multiset(ctx.hash_join(R, S, left_outer=True, radix_bits=2), want)
    assert same_multiset(ctx.nlj_join(R, S, left_outer=True), want)
    e = np.empty((0, 2), np.int32)
    assert same_multiset(ctx.hash_join(e, S, left_outer=True), _outer_want(e, S, e))
    assert same_multiset(ctx.nlj_join(e, S, left_outer=True), _outer_want(e, S, e))
    for fn in (lambda r, **k: ctx.hash_join(r, S, radix_bits=2, **k), lambda r, **k: ctx.nlj_join(r, S, **k)):
        d1, d2 = fn(R, left_outer=True, deterministic=True), fn(R, left_outer=True, deterministic=True)
        assert np.array_equal(d1, d2) and same_multiset(d1, want)
        assert same_multiset(fn(e, left_outer=True, deterministic=True), _outer_want(e, S, e))


@pytest.mark.parametrize("kind", ["hash", "hash2", "nlj"])
def test_left_outer_deterministic_order(ctx, kind):
    """Deterministic left outer join: inner rows in count/write order, then one
    null row per unmatched S tuple in S order; identical on every run; with
    RESULT16 rows, count-only and a too small buffer (nothing written)."""
    R, S = gen.c3(seed=61, nR=40000, nS=70000, key_range=90000)
    kw = {"hash": {"radix_bits": 7}, "hash2": {"radix_bits": 12, "passes": 2}, "nlj": {}}[kind]
    fn = ctx.nlj_join if kind == "nlj" else ctx.hash_join
    want = _outer_want(R, S, oracle.hash_join(R, S, nparts=1 << 10))
    d1 = fn(R, S, left_outer=True, deterministic=True, **kw)
    d2 = fn(R, S, left_outer=True, deterministic=True, **kw)
    assert np.array_equal(d1, d2) and same_multiset(d1, want)
    nul = d1[:, 0] == _abi.EO_NULL_PAYLOAD
    first_null = int(np.argmax(nul))
    assert nul[first_null:].all() and not nul[:first_null].any()  # null rows last
    # their S payloads in S order (hash: the stable partition's order, by key mod 2^bits)
    Sp = S if kind == "nlj" else S[np.argsort(S[:, 0] & ((1 << kw["radix_bits"]) - 1), kind="stable")]
    assert np.array_equal(d1[nul, 1], Sp[~np.isin(Sp[:, 0], R[:, 0]), 1])
    assert fn(R, S, left_outer=True, deterministic=True, count_only=True, **kw) == len(want)
    w16 = fn(R, S, left_outer=True, deterministic=True, result16=True, **kw)
    assert np.array_equal(w16[:, [0, 2]], d1)
    out = np.full((len(want) - 1, 2), 7, np.int32)
    import ctypes
    n = ctypes.c_uint64()
    o = ctx.opts(left_outer=True, deterministic=True, **kw)
    lib_fn = ctx.lib.eo_nlj_join if kind == "nlj" else ctx.lib.eo_hash_join
    st = lib_fn(ctx.handle, R.ctypes.data, len(R), S.ctypes.data, len(S), out.ctypes.data, len(out),
                ctypes.byref(n), ctypes.byref(o))
    assert st == _abi.EO_ECAPACITY and n.value == len(want) and (out == 7).all()


def test_left_outer_nlj_and_index_joins(ctx):
    R, S = gen.c3(seed=98, nR=30000, nS=50000, key_range=60000)
    want = _outer_want(R, S, oracle.hash_join(R, S, nparts=1 << 10))
    assert same_multiset(ctx.nlj_join(R, S, left_outer=True), want)
    for ix in (ctx.hidx_build(R), ctx.sidx_build(R), ctx.sidx_build(R, learned=True)):
        assert same_multiset(ix.join(S, left_outer=True), want)
        assert same_multiset(ix.join(S, left_outer=True, deterministic=True), want)
        assert ix.join(S, left_outer=True, count_only=True) == len(want)
        ix.free()
    ix = ctx.sidx_build(np.empty((0, 2), np.int32))
    assert same_multiset(ix.join(S, left_outer=True), _outer_want(np.empty((0, 2), np.int32), S,
                                                                  np.empty((0, 2), np.int32)))
    ix.free()


@pytest.mark.parametrize("nR,nS,dist", [(3000, 5000, 1.0), (1025, 1023, 0.5), (4100, 2100, 2.0)])
def test_nlj_dist_join_equals_oracle(ctx, nR, nS, dist):
    rng = np.random.default_rng(nR + nS)
    R = np.zeros(nR, dtype=oracle.POINT3)
    S = np.zeros(nS, dtype=oracle.POINT3)
    R["id"], S["id"] = np.arange(nR), np.arange(nS) + 10 ** 6
    R["v"] = rng.uniform(0, 12, size=(nR, 3))
    S["v"] = rng.uniform(0, 12, size=(nS, 3))
    # boundary pairs: S points at distance exactly dist, and one ulp inside, of R points
    k = min(200, nR, nS // 2)
    S["v"][:k] = R["v"][:k] + np.array([dist, 0.0, 0.0])
    S["v"][k:2 * k] = R["v"][:k] + np.array([0.0, np.nextafter(dist, 0.0), 0.0])
    want = oracle.nlj_dist_join(R, S, dist)
    got = ctx.nlj_dist_join(R, S, dist)
    assert len(got) == len(want) > 0
    assert same_multiset(got, want)
    assert ctx.nlj_dist_join(R, S, dist, count_only=True) == len(want)


@pytest.mark.parametrize("hot", [0.3, 0.7, 1.0])
def test_two_pass_hot_key(ctx, hot):
    """One key holding up to all of S: one first-pass segment far larger than the
    rest, handed out to the persistent second pass in many groups."""
    rng = np.random.default_rng(int(hot * 10))
    nR, nS = 1 << 20, 1 << 21
    R = np.stack([rng.permutation(nR), np.arange(nR)], 1).astype(np.int32)
    keys = rng.integers(0, nR, nS)
    keys[rng.random(nS) < hot] = 777
    S = np.stack([keys, np.arange(nS)], 1).astype(np.int32)
    got = ctx.hash_join(R, S, radix_bits=13, passes=2)
    assert len(got) == nS
    # every S tuple matches the single R tuple with its key (R payload = row)
    assert np.array_equal(np.sort(got[:, 1]), np.arange(nS))
    assert np.array_equal(R[got[:, 0], 0], S[got[:, 1], 0])


@pytest.mark.parametrize("nR,nS,key_range", [(1 << 20, 5000, 1 << 20), (300000, 20000, 1 << 14), (200000, 199999, 1 << 16),
                                           (1 << 22, 1 << 18, 1 << 22)])
def test_hash_join_builds_on_the_smaller_side(ctx, nR, nS, key_range):
    """An inner join with |S| < |R| builds its tables on S and streams R past
    them; the rows still come out as (R payload, S payload), in every row
    layout, and equal the oracle (also count only and on capacity overflow)."""
    R, S = gen.uniform(seed=nR % 977 + nS, nR=nR, nS=nS, key_range=key_range)
    want = oracle.hash_join(R, S, nparts=1 << 12)
    st = eo.JoinStats()
    got = ctx.hash_join(R, S, stats=st)
    assert same_multiset(got, want)
    # the plan follows the smaller relation (a side of one LDS table is not partitioned)
    if nS <= 8192:
        assert st.radix_bits == 0
    wide = ctx.hash_join(R, S, result16=True)
    assert len(wide) == len(want)
    assert same_multiset(wide[:, [0, 2]], want)
    assert np.array_equal(wide[:, 1], wide[:, 3])
    assert np.array_equal(R[wide[:, 0], 0], wide[:, 1]) and np.array_equal(S[wide[:, 2], 0], wide[:, 3])
    assert ctx.hash_join(R, S, count_only=True) == len(want)
    # a deterministic or left outer join keeps R as the build side: same rows
    assert same_multiset(ctx.hash_join(R, S, deterministic=True), want)
    dR, dS = eo.DeviceArray.from_host(ctx, R), eo.DeviceArray.from_host(ctx, S)
    small = eo.DeviceArray(ctx, 8 * 100)
    n, status = ctx.hash_join_device(dR, nR, dS, nS, small, 100)
    assert status == eo._abi.EO_ECAPACITY and n == len(want)
    rows = small.to_host(np.int32, (100, 2))
    assert np.array_equal(R[rows[:, 0], 0], S[rows[:, 1], 0])
