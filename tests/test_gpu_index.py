"""GPU parity of the resident indexes (hash, sorted, learned) through the C ABI.

Oracles: joins/CPU_JOIN/inner_cpu (hash index) and joins/CPU_JOIN/sortIndex_cpu
(sorted index) restated in oracle/cpu_join.c; the learned model is checked probe
by probe against the plain binary search of the same index.
"""
import numpy as np
import pytest

import oracle
from eating_one_amd import _abi, gen
from common import brute_join, join_count_closed_form, same_multiset

pytestmark = pytest.mark.gpu


def test_hidx_c3_full(ctx):
    R, S = gen.c3()
    ix = ctx.hidx_build(R)
    try:
        got = ix.join(S)
        assert len(got) == join_count_closed_form(R, S)
        assert same_multiset(got, oracle.hidx_join(R, S, nbuckets=1 << 16))
        assert ix.join(S, count_only=True) == len(got)
    finally:
        ix.free()


@pytest.mark.parametrize("learned", [False, True])
def test_sidx_c3_full(ctx, learned):
    R, S = gen.c3()
    ix = ctx.sidx_build(R, learned=learned)
    try:
        got = ix.join(S)
        assert same_multiset(got, oracle.sidx_join(R, S))
        keys, pays = ix.export()
        ref = oracle.sidx_build(R)
        assert np.array_equal(keys, ref[:, 0])                  # sorted keys exactly
        assert np.array_equal(np.sort(R[ref[:, 1], 1]), np.sort(pays))
        # per key, the payload multiset matches the oracle's {val, adr} index
        o1 = np.lexsort((pays, keys))
        o2 = np.lexsort((R[ref[:, 1], 1], ref[:, 0]))
        assert np.array_equal(pays[o1], R[ref[:, 1], 1][o2])
    finally:
        ix.free()


def test_index_edge_cases(ctx):
    e = np.empty((0, 2), dtype=np.int32)
    one = np.array([[5, 9]], dtype=np.int32)
    R = np.array([[-(1 << 31), 0], [(1 << 31) - 1, 1], [0, 2], [-1, 3], [-1, 4]], dtype=np.int32)
    S = np.array([[(1 << 31) - 1, 0], [-(1 << 31), 1], [-1, 2], [7, 3]], dtype=np.int32)
    for build in (lambda r: ctx.hidx_build(r), lambda r: ctx.sidx_build(r), lambda r: ctx.sidx_build(r, learned=True)):
        ix = build(e)
        assert len(ix.join(one)) == 0
        ix.free()
        ix = build(one)
        assert len(ix.join(e)) == 0
        assert same_multiset(ix.join(one), np.array([[9, 9]]))
        ix.free()
        ix = build(R)
        assert same_multiset(ix.join(S), brute_join(R, S))
        ix.free()


@pytest.mark.parametrize("seed,nR,nS,kr", [(1, 5000, 7000, 300), (2, 100000, 50000, 1 << 20), (3, 70001, 3, 10)])
def test_indexes_random(ctx, seed, nR, nS, kr):
    R, S = gen.uniform(seed, nR, nS, kr, lo=-(kr // 2))
    want = brute_join(R, S)
    for ix in (ctx.hidx_build(R), ctx.sidx_build(R), ctx.sidx_build(R, learned=True)):
        try:
            assert same_multiset(ix.join(S), want)
        finally:
            ix.free()


def _check_learned(ctx, R, S):
    ix = ctx.sidx_build(R, learned=True)
    try:
        info = ix.info()
        assert info["learned"] and info["leaves"] == 4096
        keys, _ = ix.export()
        lb_bin = ix.lookup(S, _abi.EO_LOOKUP_BINARY)
        lb_lrn = ix.lookup(S, _abi.EO_LOOKUP_LEARNED)
        want = np.searchsorted(keys, S[:, 0], side="left").astype(np.uint64)
        assert np.array_equal(lb_bin, want), "binary search differs from numpy's lower bound"
        present = np.isin(S[:, 0], keys)
        assert np.array_equal(lb_lrn[present], lb_bin[present]), "learned position differs from binary search"
        # absent keys: the learned window holds no equal key
        pos = np.minimum(lb_lrn[~present], len(keys) - 1).astype(np.int64)
        assert not np.any(keys[pos] == S[~present, 0])
        assert same_multiset(ix.join(S), ix.join(S, binary_search=True))
        return info
    finally:
        ix.free()


def test_learned_index_c4_shape(ctx):
    R, S = gen.c4(nR=1 << 21, nS=1 << 22)
    info = _check_learned(ctx, R, S)
    assert info["max_err"] < 1 << 16


def test_learned_index_duplicates_and_uniform(ctx):
    R, S = gen.uniform(5, 1 << 20, 1 << 20, 1 << 18)
    _check_learned(ctx, R, S)


@pytest.mark.parametrize("kind", ["hash", "sorted", "learned"])
def test_index_join_fused_and_deterministic(ctx, kind):
    """Default one-pass index join (atomic output reservation) vs count/scan/write."""
    import ctypes
    from eating_one_amd import _abi
    R, S = gen.c3(seed=77, nR=50000, nS=80000, key_range=20000)
    ix = ctx.hidx_build(R) if kind == "hash" else ctx.sidx_build(R, learned=(kind == "learned"))
    want = oracle.hash_join(R, S, nparts=1 << 10)
    fused = ix.join(S)
    d1 = ix.join(S, deterministic=True)
    d2 = ix.join(S, deterministic=True)
    assert same_multiset(fused, want) and same_multiset(d1, want)
    assert np.array_equal(d1, d2)
    assert ix.join(S, count_only=True) == len(want)
    out = np.empty((len(want) - 5, 2), dtype=np.int32)
    n = ctypes.c_uint64()
    o = ctx.opts()
    st = ctx.lib.eo_index_join(ctx.handle, ix.handle, S.ctypes.data, len(S), out.ctypes.data, len(want) - 5,
                               ctypes.byref(n), ctypes.byref(o))
    assert st == _abi.EO_ECAPACITY and n.value == len(want)
    ix.free()


def test_sidx_import_from_reference_index_file(ctx, tmp_path):
    """A BUCKET index file read from disk gives the same index as building it on the device."""
    import eating_one_amd as eo
    R, S = gen.c3(seed=81, nR=70000, nS=90000, key_range=30000)
    bucket = R[np.argsort(R[:, 0], kind="stable")]
    path = tmp_path / "index"
    eo.write_ref_index(str(path), bucket)
    for learned in (False, True):
        imp = ctx.sidx_import(eo.read_ref_index(str(path)), learned=learned)
        built = ctx.sidx_build(R, learned=learned)
        k1, p1 = imp.export()
        k2, p2 = built.export()
        assert np.array_equal(k1, k2) and np.array_equal(k1, bucket[:, 0])
        assert same_multiset(imp.join(S), oracle.hash_join(R, S, nparts=1 << 10))
        if learned:
            assert imp.info()["max_err"] == built.info()["max_err"]
        imp.free()
        built.free()
    with pytest.raises(eo.EOError):
        ctx.sidx_import(R)  # not sorted
