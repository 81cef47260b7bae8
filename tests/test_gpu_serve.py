"""GPU tests of the device generator and the request loop (continuous batching).

Each served result must equal the oracle's result for that request alone, in
both serving modes, whatever batch it landed in.
"""
import numpy as np
import pytest

import eating_one_amd as eo
import oracle
from eating_one_amd import DeviceArray, gen
from common import same_multiset

pytestmark = pytest.mark.gpu


def test_gen_build_is_a_seeded_permutation(ctx):
    for n in (1, 2, 1000, 100003, 1 << 20):
        d = DeviceArray(ctx, 8 * n)
        ctx.gen_build(d, n, seed=12345)
        t = d.to_host(np.int32, (n, 2))
        assert np.array_equal(np.sort(t[:, 0]), np.arange(n, dtype=np.int32)), n
        assert np.array_equal(t[:, 1], np.arange(n, dtype=np.int32))
        if n >= 1000:
            assert not np.array_equal(t[:, 0], np.arange(n, dtype=np.int32))
            d2 = DeviceArray(ctx, 8 * n)
            ctx.gen_build(d2, n, seed=12345)
            assert np.array_equal(d2.to_host(np.int32, (n, 2)), t)
            ctx.gen_build(d2, n, seed=777)
            assert not np.array_equal(d2.to_host(np.int32, (n, 2)), t)


@pytest.mark.parametrize("sel", [0.0, 0.1, 0.55, 1.0])
def test_gen_probe_selectivity_and_join_count(ctx, sel):
    nR, nS = 50000, 200000
    dR, dS = DeviceArray(ctx, 8 * nR), DeviceArray(ctx, 8 * nS)
    ctx.gen_build(dR, nR, seed=1)
    ctx.gen_probe(dS, nS, nR, sel, seed=2)
    R = dR.to_host(np.int32, (nR, 2))
    S = dS.to_host(np.int32, (nS, 2))
    hits = int((S[:, 0] < nR).sum())
    assert (S[:, 0] >= 0).all() and (S[:, 0] < (1 << 31) - 1).all()
    assert np.array_equal(S[:, 1], np.arange(nS, dtype=np.int32))
    assert abs(hits / nS - sel) < 0.01
    # every hit matches exactly one build tuple
    n, st = ctx.hash_join_device(dR, nR, dS, nS, None, 0)
    assert n == hits
    assert len(oracle.hash_join(R, S, nparts=1024)) == hits


def _stream(ctx, seed, n_req, lo, hi, n_tables=3, reuse=0.25):
    rng = np.random.default_rng(seed)
    tables = []
    for k in range(n_tables):
        nR = int(2 ** rng.uniform(lo, hi))
        R, _ = gen.c2(seed=seed * 10 + k, nR=nR, nS=1)
        tables.append((R, ctx.table_create(R)))
    items = []
    for i in range(n_req):
        if rng.random() < reuse:
            R, t = tables[int(rng.integers(0, n_tables))]
            dR = None
        else:
            nR = int(2 ** rng.uniform(lo, hi))
            R, _ = gen.c2(seed=seed * 1000 + i, nR=nR, nS=1)
            t = None
            dR = DeviceArray.from_host(ctx, R)
        nS = int(2 ** rng.uniform(lo, hi))
        sel = float(rng.uniform(0.1, 1.0))
        keys = np.where(rng.random(nS) < sel, rng.integers(0, len(R), nS), rng.integers(len(R), 1 << 30, nS))
        S = np.stack([keys.astype(np.int32), np.arange(nS, dtype=np.int32)], 1)
        want = oracle.hash_join(R, S, nparts=max(1, len(R) // 64))
        dS = DeviceArray.from_host(ctx, S)
        dout = DeviceArray(ctx, max(len(want), 1) * 8)
        items.append(dict(R=R, S=S, want=want, dR=dR, dS=dS, dout=dout, table=t))
    reqs = [eo.request(R=it["dR"], nR=len(it["R"]), S=it["dS"], nS=len(it["S"]), out=it["dout"],
                       out_cap=len(it["want"]), table=it["table"], tag=i) for i, it in enumerate(items)]
    return tables, items, reqs


@pytest.mark.parametrize("mode", [eo._abi.EO_SERVE_BATCHED, eo._abi.EO_SERVE_PER_JOIN])
def test_serve_results_equal_oracle(ctx, mode):
    tables, items, reqs = _stream(ctx, 31, 40, 8, 18)
    rng = np.random.default_rng(5)
    arrival = np.cumsum(rng.exponential(0.2, size=len(reqs)))  # 5000 requests/s
    res, lat, st = ctx.serve(reqs, arrival, mode=mode)
    for i, ((tag, n, status), it) in enumerate(zip(res, items)):
        assert tag == i and status == 0 and n == len(it["want"]), i
        got = it["dout"].to_host(np.int32, (n, 2)) if n else np.empty((0, 2), np.int32)
        assert same_multiset(got, it["want"]), i
    assert (lat > 0).all()
    assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["n_out"] == sum(len(it["want"]) for it in items)
    assert st["tuples"] == sum(len(it["R"]) + len(it["S"]) for it in items)
    if mode == eo._abi.EO_SERVE_BATCHED:
        assert st["batches"] <= len(reqs)
    for _, t in tables:
        t.free()


def test_serve_budget_admission(ctx):
    """A small budget forces many batches, exclusive runs and a refusal."""
    tables, items, reqs = _stream(ctx, 41, 24, 10, 17, reuse=0.0)
    sizes = [8 * (len(it["R"]) + 2 * len(it["S"])) for it in items]
    budget = int(np.median(sizes) * 6)
    huge = int(np.argmax(sizes))
    res, lat, st = ctx.serve(reqs, None, hbm_budget=budget, max_batch=4)
    assert st["peak_admitted"] <= budget
    assert st["batches"] >= len(reqs) // 4
    for i, ((tag, n, status), it) in enumerate(zip(res, items)):
        if status == eo._abi.EO_EBUDGET:
            continue
        assert status == 0 and n == len(it["want"]), i
        got = it["dout"].to_host(np.int32, (n, 2)) if n else np.empty((0, 2), np.int32)
        assert same_multiset(got, it["want"]), i
    # a budget below the largest request refuses exactly the requests above it
    res2, _, _ = ctx.serve(reqs, None, hbm_budget=sizes[huge] // 3)
    assert res2[huge][2] == eo._abi.EO_EBUDGET
    assert any(r[2] == 0 for r in res2)
    for _, t in tables:
        t.free()


@pytest.mark.parametrize("op", ["hash", "nlj", "hidx", "sidx", "learned", "stream-hash", "stream-nlj"])
def test_cli_counts(op):
    """eo_join <op> <|S|> <|R|>: with selectivity 1 every S key matches exactly one R key."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "eating_one_amd", "eo_join")
    nS, nR = 300000, 70000
    r = subprocess.run([exe, op, str(nS), str(nR), "7", "1.0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    count = int([l for l in lines if l.startswith("count:")][0].split()[1])
    assert count == nS
    assert sum(l.startswith("pair ") for l in lines) == 3
    r = subprocess.run([exe, op, str(nS), str(nR), "7", "0.0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "count: 0" in r.stdout


def test_serve_deadlines_and_cancellation(ctx):
    """Requests cancelled before admission, or waiting past their admission
    deadline, complete without running (output untouched); the others equal
    the oracle.  One slot and batches of one make later requests wait."""
    tables, items, reqs = _stream(ctx, 51, 24, 12, 18, reuse=0.0)
    n = len(reqs)
    sentinel = np.full((1, 2), -7, np.int32)
    for it in items:  # a sentinel row where each output starts
        if len(it["want"]):
            it["dout"].ctx.lib.eo_copy_h2d(ctx.handle, it["dout"].ptr, sentinel.ctypes.data, 8)
    cancel = np.zeros(n, np.uint32)
    cancel[[3, 4, 17]] = 1
    deadline = np.zeros(n)
    deadline[1::2] = 1e-3  # odd requests must start within a microsecond of arriving
    res, lat, st = ctx.serve(reqs, None, slots=1, max_batch=1, deadline_ms=deadline, cancel=cancel)
    ran = 0
    for i, ((tag, cnt, status), it) in enumerate(zip(res, items)):
        assert tag == i
        if cancel[i]:
            assert status == eo._abi.EO_ECANCELED and cnt == 0, i
        elif status == eo._abi.EO_ETIMEDOUT:
            assert deadline[i] > 0 and cnt == 0, i
        else:
            assert status == 0 and cnt == len(it["want"]), i
            got = it["dout"].to_host(np.int32, (cnt, 2)) if cnt else np.empty((0, 2), np.int32)
            assert same_multiset(got, it["want"]), i
            ran += 1
            continue
        if len(it["want"]):  # did not run: the output is untouched
            assert np.array_equal(it["dout"].to_host(np.int32, (1, 2)), sentinel), i
    assert st["cancelled"] == 3
    # all requests arrive at once and run one at a time: most odd ones expire
    assert st["expired"] >= 5
    assert ran == n - st["cancelled"] - st["expired"]
    assert st["refused"] == 0
    # no deadlines and no flags: everything runs
    res2, _, st2 = ctx.serve(reqs, None, slots=1, max_batch=1)
    assert all(r[2] == 0 and r[1] == len(it["want"]) for r, it in zip(res2, items))
    assert st2["expired"] == st2["cancelled"] == 0
    for _, t in tables:
        t.free()


def test_serve_cancel_from_another_thread(ctx):
    """Flags set while the loop runs stop the requests not yet admitted."""
    import threading
    import time
    tables, items, reqs = _stream(ctx, 53, 16, 12, 16, reuse=0.0)
    n = len(reqs)
    cancel = np.zeros(n, np.uint32)
    arrival = np.linspace(0.0, 400.0, n)  # one request every ~27 ms

    def stop_later():
        time.sleep(0.1)
        cancel[:] = 1

    th = threading.Thread(target=stop_later)
    th.start()
    res, lat, st = ctx.serve(reqs, arrival, cancel=cancel)
    th.join()
    statuses = [r[2] for r in res]
    assert statuses[0] == 0 and statuses[-1] == eo._abi.EO_ECANCELED
    first_cancelled = statuses.index(eo._abi.EO_ECANCELED)
    assert all(s == 0 for s in statuses[:first_cancelled])
    assert all(s == eo._abi.EO_ECANCELED for s in statuses[first_cancelled:])
    for (tag, cnt, status), it in zip(res[:first_cancelled], items):
        assert cnt == len(it["want"])
    assert st["cancelled"] == n - first_cancelled
    assert st["wall_ms"] < 350  # the loop ended at the cancellation, not at the last arrival
    for _, t in tables:
        t.free()


def _defeating_probe(rng, n, key_range):
    """Probe keys that defeat the partition sample: every line the sample
    reads (one 128-byte line in each group of 16, offsets hashed from tile and
    group) holds spread keys, every other line the key 5."""
    keys = np.full(n, 5, dtype=np.int32)
    for t in range(n // 16384):
        for j in range(64):
            h = ((t * 64 + j) * 0x9E3779B1) & 0xFFFFFFFF
            h ^= h >> 15
            h = (h * 0x85EBCA77) & 0xFFFFFFFF
            h ^= h >> 13
            b = t * 16384 + (j * 16 + (h >> 28)) * 16
            keys[b:b + 16] = rng.integers(0, key_range, size=16, dtype=np.int32)
    return np.stack([keys, np.arange(n, dtype=np.int32)], 1)


def test_serve_redoes_an_overflowing_sampled_batch(ctx):
    """A served batch of 256+ requests samples its large relations; one
    request's probe side defeats the sample, the batch is redone with exact
    partitions inside the loop, and every result still equals the oracle."""
    rng = np.random.default_rng(0x5F)
    R = np.stack([rng.permutation(1 << 20).astype(np.int32), np.arange(1 << 20, dtype=np.int32)], 1)
    S = _defeating_probe(rng, 1 << 21, 1 << 20)
    want0 = oracle.hash_join(R, S, nparts=1 << 12)
    dR, dS = DeviceArray.from_host(ctx, R), DeviceArray.from_host(ctx, S)
    dout = DeviceArray(ctx, len(want0) * 8)
    tables, items, reqs = _stream(ctx, 51, 259, 6, 11, n_tables=2)
    reqs = [eo.request(R=dR, nR=len(R), S=dS, nS=len(S), out=dout, out_cap=len(want0), tag=1000)] + reqs
    res, lat, st = ctx.serve(reqs, None, mode=eo._abi.EO_SERVE_BATCHED)
    tag, n, status = res[0]
    assert tag == 1000 and status == 0 and n == len(want0)
    assert same_multiset(dout.to_host(np.int32, (n, 2)), want0)
    for i, ((tag, n, status), it) in enumerate(zip(res[1:], items)):
        assert status == 0 and n == len(it["want"]), i
        got = it["dout"].to_host(np.int32, (n, 2)) if n else np.empty((0, 2), np.int32)
        assert same_multiset(got, it["want"]), i
    for _, t in tables:
        t.free()
