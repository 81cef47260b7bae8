"""Parity at the full benchmark sizes through size-independent properties.

The oracle cannot run 2^27 x 2^28 in seconds; at that size the checks are
properties every correct result has and a plausible mistake breaks:
  C2 (R a permutation of [0, 2^27), every S key drawn from R): the count is
  |S|, every S payload appears exactly once, and the R and S keys of every
  pair are equal (a swapped field, a dropped or duplicated tuple, or a wrong
  partition would each fail one of them);
  C4: the learned model's lower bound equals binary search for every one of
  the 2^27 probes whose key is in R (and finds no equal key for the others).
"""
import numpy as np
import pytest

import eating_one_amd as eo
from eating_one_amd import _abi, gen

pytestmark = pytest.mark.gpu


def test_c2_full_size_properties(ctx):
    nR, nS = 1 << 27, 1 << 28
    R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS)
    dR, dS = eo.DeviceArray.from_host(ctx, R), eo.DeviceArray.from_host(ctx, S)
    dout = eo.DeviceArray(ctx, nS * 8)
    # the default plan (14 bits, two passes) sizes its partitions from a sample;
    # EO_F_EXACT_PARTITIONS takes the histogram pass: both give every row
    for flags, sampled in ((0, 1), (_abi.EO_F_EXACT_PARTITIONS, 0)):
        js = eo.JoinStats()
        n, st = ctx.hash_join_device(dR, nR, dS, nS, dout, nS, flags=flags, stats=js)
        assert st == 0 and n == nS
        assert js.passes == 2 and js.sampled == sampled
        pairs = dout.to_host(np.int32, (nS, 2))
        assert np.array_equal(np.bincount(pairs[:, 1], minlength=nS), np.ones(nS, dtype=np.int64))
        assert np.array_equal(R[pairs[:, 0], 0], S[pairs[:, 1], 0])
    dout.close()


def test_c4_full_size_learned_equals_binary(ctx):
    """For every probe whose key is in R the learned lower bound equals binary
    search; for an absent key the learned window holds no equal key."""
    R, S = gen.c4(nR=1 << 26, nS=1 << 27)
    ix = ctx.sidx_build(R, learned=True)
    try:
        lb = ix.lookup(S, method=_abi.EO_LOOKUP_BINARY)
        ll = ix.lookup(S, method=_abi.EO_LOOKUP_LEARNED)
        keys, _ = ix.export()
    finally:
        ix.free()
    n = len(keys)
    hit = lb < n
    present = np.zeros(len(S), bool)
    present[hit] = keys[lb[hit]] == S[hit, 0]
    assert int(present.sum()) == int(round(0.8 * len(S)))        # 80 % of the probes are R keys
    assert np.array_equal(ll[present], lb[present])
    pos = np.minimum(ll[~present], n - 1).astype(np.int64)
    assert not np.any(keys[pos] == S[~present, 0])


def test_c2_full_size_streamed_from_host(ctx):
    """The C2 relations held in pinned host memory and joined by eo_stream_join
    (S in 16 chunks of 2^24): the same properties as the in-HBM join."""
    nR, nS = 1 << 27, 1 << 28
    R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS)
    hR, hS, hO = eo.HostArray(ctx, R.nbytes), eo.HostArray(ctx, S.nbytes), eo.HostArray(ctx, 8 * nS)
    try:
        vR, vS, vO = hR.view(np.int32, R.shape), hS.view(np.int32, S.shape), hO.view(np.int32, (nS, 2))
        vR[:] = R
        vS[:] = S
        st = _abi.StreamStats()
        n, status = ctx.stream_join(vR, vS, out=vO, stats=st)
        assert status == _abi.EO_OK and n == nS and st.pinned == 7 and st.s_chunks == 16
        pairs = vO.copy()
    finally:
        hO.close()
        hS.close()
        hR.close()
    assert np.array_equal(np.bincount(pairs[:, 1], minlength=nS), np.ones(nS, dtype=np.int64))
    assert np.array_equal(R[pairs[:, 0], 0], S[pairs[:, 1], 0])


def test_two_pass_half_tables_outer_and_deterministic(ctx):
    """|R| = 2^26 needs two partition passes, so the fused join takes half LDS
    tables (13 bits); selectivity 0.5 (absent S keys are >= |R|).  Properties:
    inner rows = exactly the present S tuples once each with equal keys; the
    left outer join adds one null row per absent S tuple; the deterministic
    join (ordered partition, fixed table order) gives the same rows in the
    same order twice, and the same set as the fused join."""
    nR, nS = 1 << 26, 1 << 27
    R, S = gen.c2(seed=0xE01C5, nR=nR, nS=nS, selectivity=0.5)
    present = S[:, 0] < nR
    dR, dS = eo.DeviceArray.from_host(ctx, R), eo.DeviceArray.from_host(ctx, S)
    dout = eo.DeviceArray(ctx, nS * 8)
    try:
        st = eo.JoinStats()
        n, status = ctx.hash_join_device(dR, nR, dS, nS, dout, nS, stats=st)
        assert status == 0 and n == int(present.sum())
        assert st.lds_tuples == 8192 and st.passes == 2 and st.radix_bits == 13 and st.sampled == 1
        pairs = dout.to_host(np.int32, (n, 2))
        cnt = np.bincount(pairs[:, 1], minlength=nS)
        assert cnt.max() == 1 and np.array_equal(cnt.astype(bool), present)
        assert np.array_equal(R[pairs[:, 0], 0], S[pairs[:, 1], 0])

        n2, status = ctx.hash_join_device(dR, nR, dS, nS, dout, nS, left_outer=True)
        assert status == 0 and n2 == nS
        rows = dout.to_host(np.int32, (nS, 2))
        null = rows[:, 0] == _abi.EO_NULL_PAYLOAD
        assert np.array_equal(np.bincount(rows[:, 1], minlength=nS), np.ones(nS, dtype=np.int64))
        assert np.array_equal(np.sort(rows[null, 1]), np.flatnonzero(~present))
        assert np.array_equal(R[rows[~null, 0], 0], S[rows[~null, 1], 0])

        runs = []
        for _ in range(2):
            st = eo.JoinStats()
            n3, status = ctx.hash_join_device(dR, nR, dS, nS, dout, nS, deterministic=True, stats=st)
            assert status == 0 and n3 == n and st.lds_tuples == 8192
            runs.append(dout.to_host(np.int32, (n, 2)))
        assert np.array_equal(runs[0], runs[1])
        a = runs[0][np.argsort(runs[0][:, 1], kind="stable")]
        b = pairs[np.argsort(pairs[:, 1], kind="stable")]
        assert np.array_equal(a, b)
    finally:
        for x in (dout, dS, dR):
            x.close()


@pytest.mark.parametrize("theta,sel,bits,passes", [(0.9, 0.1, 0, 0), (0.5, 1.0, 0, 0), (0.0, 1.0, 11, 1)],
                         ids=["zipf0.9-sel0.1", "zipf0.5-sel1.0", "one-pass-11-bits"])
def test_c2_full_size_variants(ctx, theta, sel, bits, passes):
    """The other C2 shapes at 2^27 x 2^28: Zipf-skewed probe keys (hot keys
    land in a few partitions, which the planner splits into span tiles),
    10 % selectivity, and a forced one-pass 11-bit partition.  Properties:
    count = the S keys present in R, each present S payload exactly once,
    equal keys in every pair; the deterministic join repeats its rows in the
    same order and equals the fused join as a set."""
    nR, nS = 1 << 27, 1 << 28
    print(f"generating C2 theta={theta} selectivity={sel}", flush=True)
    R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS, theta=theta, selectivity=sel)
    present = S[:, 0] < nR
    n_present = int(present.sum())
    dR, dS = eo.DeviceArray.from_host(ctx, R), eo.DeviceArray.from_host(ctx, S)
    dout = eo.DeviceArray(ctx, n_present * 8)
    try:
        print(f"fused join, radix_bits={bits} passes={passes}", flush=True)
        st = eo.JoinStats()
        n, status = ctx.hash_join_device(dR, nR, dS, nS, dout, n_present, radix_bits=bits, passes=passes, stats=st)
        assert status == 0 and n == n_present
        if bits:
            assert st.radix_bits == bits and st.passes == passes
        # two-pass plans size their partitions from a sample (skewed keys included)
        assert st.sampled == (1 if st.passes == 2 else 0)
        pairs = dout.to_host(np.int32, (n, 2))
        cnt = np.bincount(pairs[:, 1], minlength=nS)
        assert cnt.max() == 1 and np.array_equal(cnt.astype(bool), present)
        assert np.array_equal(R[pairs[:, 0], 0], S[pairs[:, 1], 0])
        print("deterministic join, twice", flush=True)
        runs = []
        for _ in range(2):
            n2, status = ctx.hash_join_device(dR, nR, dS, nS, dout, n_present, radix_bits=bits, passes=passes,
                                              deterministic=True)
            assert status == 0 and n2 == n
            runs.append(dout.to_host(np.int32, (n, 2)))
        assert np.array_equal(runs[0], runs[1])
        by_s = lambda p: p[np.argsort(p[:, 1])]
        assert np.array_equal(by_s(runs[0]), by_s(pairs))
    finally:
        for x in (dout, dS, dR):
            x.close()
