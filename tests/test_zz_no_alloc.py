"""eo_hash_join_noalloc: the hash join with its scratch in one fixed device
buffer (no allocation in the call).  Its results must equal eo_hash_join's on
the same inputs (the same match count and the same rows as a multiset) on
two calls in a row, the second smaller than the first (the buffer keeps the
first call's contents), and it must refuse what it cannot do."""
import ctypes

import numpy as np
import pytest

import eating_one_amd as eo
from eating_one_amd import DeviceArray, _abi, gen
from common import same_multiset

pytestmark = pytest.mark.gpu


def _noalloc(ctx, dR, nR, dS, nS, dout, cap, **kw):
    n = ctypes.c_uint64()
    o = ctx.opts(where=_abi.EO_DEVICE, **kw)
    st = ctx.lib.eo_hash_join_noalloc(ctx.handle, dR.ptr if dR else None, nR, dS.ptr if dS else None, nS,
                                      dout.ptr if dout else None, cap, ctypes.byref(n), ctypes.byref(o))
    return n.value, st


def _rows(d, n, cols=2):
    return d.to_host(np.int32, (n, cols)) if n else np.empty((0, cols), np.int32)


@pytest.mark.parametrize("flags", [{}, {"left_outer": True}, {"result16": True}])
def test_noalloc_equals_hash_join_twice(ctx, flags):
    cols = 4 if flags.get("result16") else 2
    # first call larger (two partitions' worth of scratch reused), second smaller
    for i, (nR, nS, sel) in enumerate([(1 << 20, 3 << 20, 0.7), (70000, 300001, 0.4)]):
        print(f"call {i}: |R| = {nR}, |S| = {nS}, {flags}", flush=True)
        R, S = gen.c2(seed=0x5EED + i, nR=nR, nS=nS, selectivity=sel, theta=0.5)
        dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
        cap = nS
        a, b = DeviceArray(ctx, cap * 4 * cols), DeviceArray(ctx, cap * 4 * cols)
        n_ref, st_ref = ctx.hash_join_device(dR, nR, dS, nS, a, cap, **flags)
        n, st = _noalloc(ctx, dR, nR, dS, nS, b, cap, **flags)
        assert st_ref == 0 and st == 0 and n == n_ref
        ra, rb = _rows(a, n, cols), _rows(b, n, cols)
        if cols == 2:
            assert same_multiset(ra, rb)
        else:
            assert np.array_equal(ra[np.lexsort(ra.T[::-1])], rb[np.lexsort(rb.T[::-1])])
        # count only
        n0, st0 = _noalloc(ctx, dR, nR, dS, nS, None, 0, **flags)
        assert st0 == 0 and n0 == n_ref
        for d in (a, b, dS, dR):
            d.close()


def test_noalloc_refuses_what_needs_allocation(ctx):
    R, S = gen.c2(seed=9, nR=1000, nS=2000)
    dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
    dout = DeviceArray(ctx, 8 * 2000)
    # host pointers, the deterministic order
    n = ctypes.c_uint64()
    o = ctx.opts(where=_abi.EO_HOST)
    assert ctx.lib.eo_hash_join_noalloc(ctx.handle, R.ctypes.data, 1000, S.ctypes.data, 2000, None, 0,
                                        ctypes.byref(n), ctypes.byref(o)) == _abi.EO_EINVAL
    assert _noalloc(ctx, dR, 1000, dS, 2000, dout, 2000, deterministic=True)[1] == _abi.EO_EINVAL
    # a join whose scratch passes the fixed buffer: refused with the sizes
    big = 1 << 23
    dbig = DeviceArray(ctx, 8 * big)
    ctx.gen_build(dbig, big, seed=3)
    n_big, st = _noalloc(ctx, dbig, big, dbig, big, None, 0)
    assert st == _abi.EO_EINVAL
    assert "fixed buffer" in ctx.lib.eo_last_error().decode()
    # and the entry point still works after the refusals
    n_ok, st = _noalloc(ctx, dR, 1000, dS, 2000, dout, 2000)
    assert st == 0 and n_ok == ctx.hash_join_device(dR, 1000, dS, 2000, None, 0)[0]
    for d in (dbig, dout, dS, dR):
        d.close()
