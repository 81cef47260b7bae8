"""GPU tests of the autotuner: tuned plans change speed, never results."""
import numpy as np
import pytest

import eating_one_amd as eo
import oracle
from eating_one_amd import DeviceArray, gen
from common import same_multiset

pytestmark = pytest.mark.gpu


def test_tune_then_join_equals_oracle(ctx, tmp_path):
    R, S = gen.c2(seed=0xA1, nR=1 << 17, nS=1 << 19, theta=0.5)
    dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
    want = oracle.hash_join(R, S, nparts=1 << 12)
    dout = DeviceArray(ctx, len(want) * 8)
    ctx.tune_clear()
    assert ctx.tune_lookup(len(R), len(S)) is None
    res = ctx.tune_hash_join(dR, len(R), dS, len(S), dout, len(want), reps=2)
    assert res["candidates"] >= 3 and res["ms"] > 0 and res["default_ms"] > 0
    assert res["ms"] <= res["default_ms"] * 1.5
    entry = ctx.tune_lookup(len(R), len(S))
    assert entry["radix_bits"] == res["radix_bits"] and entry["passes"] == res["passes"]
    assert entry["lds_tuples"] == res["lds_tuples"] and res["lds_tuples"] in (16384, 8192)
    # the plain call now runs the tuned plan; same multiset as the oracle
    st = eo.JoinStats()
    got = ctx.hash_join(R, S, stats=st)
    assert st.radix_bits == res["radix_bits"] and st.passes == res["passes"]
    assert st.lds_tuples == res["lds_tuples"]
    assert same_multiset(got, want)
    # every LDS table size the tuner tries gives the oracle's rows
    for lds in (8192, 16384):
        assert same_multiset(ctx.hash_join(R, S, lds_tuples=lds), want)
    # a size class that was not tuned is untouched
    assert ctx.tune_lookup(len(R) * 4, len(S)) is None
    # save / clear / load round trip
    path = str(tmp_path / "tune.txt")
    ctx.tune_save(path)
    ctx.tune_clear()
    assert ctx.tune_lookup(len(R), len(S)) is None
    ctx.tune_load(path)
    assert ctx.tune_lookup(len(R), len(S))["radix_bits"] == res["radix_bits"]
    assert ctx.tune_lookup(len(R), len(S))["lds_tuples"] == res["lds_tuples"]
    # batches use the table too
    b = ctx.batch()
    (tag, n, status), = b.run([eo.request(R=dR, nR=len(R), S=dS, nS=len(S), out=dout, out_cap=len(want))])
    assert n == len(want) and same_multiset(dout.to_host(np.int32, (n, 2)), want)
    b.free()
    ctx.tune_clear()


def test_tune_load_rejects_bad_tables(ctx, tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("not a table\n")
    with pytest.raises(eo.EOError):
        ctx.tune_load(str(p))
    p.write_text("eojoin-tune 1\n17 19 40 1 0 1.0\n")  # 40 radix bits: invalid plan
    with pytest.raises(eo.EOError):
        ctx.tune_load(str(p))
    p.write_text("eojoin-tune 1\n17 19 12 1\n")  # truncated row
    with pytest.raises(eo.EOError):
        ctx.tune_load(str(p))
    assert ctx.tune_lookup(1 << 17, 1 << 19) is None
    # a version-1 table (no LDS column) still loads
    p.write_text("eojoin-tune 1\n17 19 4 1 0 1.0\n")
    ctx.tune_load(str(p))
    e = ctx.tune_lookup(1 << 17, 1 << 19)
    assert e["radix_bits"] == 4 and e["lds_tuples"] == 0
    ctx.tune_clear()
