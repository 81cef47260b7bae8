"""Headline benchmark: 2^27 x 2^28 radix hash join on one MI355X.

One step = one eo_hash_join over device-resident R and S (inputs already in
HBM when the timed region starts; output pairs written to a device buffer).
Prints one JSON line:
  value      = (|R| + |S|) / seconds per join   (join tuples/s)
  roofline   = the dominant kernel's algorithmic bytes per launch / its mean
               launch time (HIP events on the join's stream) vs 8 TB/s
  pipeline   = whole-join algorithmic bytes 24(|R|+|S|) + 8|out| / join time
  cpu_baseline = the CPU oracle's hash join on a bounded sample, host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(sample_log2_r: int, theta: float, sel: float, threads: int):
    """CPU_JOIN hash join (oracle restatement) on a bounded sample of the same
    workload, partitions scaled to keep the reference's tuples-per-partition."""
    import oracle
    from eating_one_amd import gen

    nR, nS = 1 << sample_log2_r, 1 << (sample_log2_r + 1)
    R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS, theta=theta, selectivity=sel)
    # CPU_JOIN uses NB_BKTENT = 2^20 partitions for |R| = 2^27: 128 R tuples per partition
    nparts = max(1, nR // 128)
    t0 = time.perf_counter()
    if threads == 1:
        n = oracle.hash_join(R, S, nparts=nparts, count_only=False)
    else:
        n = oracle.hash_join_mt(R, S, nparts=nparts, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": (nR + nS) / dt,
        "unit": "tuples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU_JOIN hash join restated (oracle/cpu_join.c), |R|=2^{sample_log2_r}, "
                  f"|S|=2^{sample_log2_r + 1}, theta={theta}, selectivity={sel}, {nparts} partitions "
                  f"(128 R tuples each, as 2^20 partitions at |R|=2^27), {len(n)} pairs, {dt:.2f} s",
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2-r", type=int, default=27)
    ap.add_argument("--theta", type=float, default=0.0)
    ap.add_argument("--selectivity", type=float, default=1.0)
    ap.add_argument("--radix-bits", type=int, default=0)
    ap.add_argument("--passes", type=int, default=0)
    ap.add_argument("--lds", type=int, default=0)
    ap.add_argument("--s-tile", type=int, default=0)
    ap.add_argument("--cpu-log2-r", type=int, default=24)
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        return 2

    import numpy as np

    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    nR, nS = 1 << a.log2_r, 1 << (a.log2_r + 1)
    log(f"generating C2 workload |R|=2^{a.log2_r} |S|=2^{a.log2_r + 1} theta={a.theta} sel={a.selectivity}")
    t0 = time.perf_counter()
    R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS, theta=a.theta, selectivity=a.selectivity)
    log(f"generated in {time.perf_counter() - t0:.1f} s")

    ctx = eo.Context(0)
    info = ctx.device_info()
    log(f"device {info}")
    dR = eo.DeviceArray.from_host(ctx, R)
    dS = eo.DeviceArray.from_host(ctx, S)
    del R
    cap = nS if a.selectivity >= 1.0 else int(nS * min(1.0, a.selectivity * 1.2 + 0.01))
    n_expected = None
    kw = dict(radix_bits=a.radix_bits, passes=a.passes, lds_tuples=a.lds, s_tile=a.s_tile)
    n_out, st = ctx.hash_join_device(dR, nR, dS, nS, None, 0, **kw)
    cap = max(cap, n_out)
    dout = eo.DeviceArray(ctx, cap * 8)
    log(f"count-only pass: {n_out} pairs")

    stats = _abi.JoinStats()

    def step():
        n, st = ctx.hash_join_device(dR, nR, dS, nS, dout, cap, stats=stats, **kw)
        assert st == _abi.EO_OK, st
        return n

    for _ in range(a.warmup):
        n_expected = step()
    ctx.sync()
    phases = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        n = step()
        assert n == n_expected
        phases.append((stats.ms_partition, stats.ms_count, stats.ms_write, stats.ms_total))
    ctx.sync()
    dt = (time.perf_counter() - t0) / a.steps
    ph = np.array(phases).mean(axis=0)
    log(f"per join {dt * 1e3:.3f} ms; device phases partition {ph[0]:.3f} count {ph[1]:.3f} write {ph[2]:.3f} "
        f"total {ph[3]:.3f} ms; bits={stats.radix_bits} passes={stats.passes} tiles={stats.tiles} "
        f"lds={stats.lds_tuples}")

    alg_bytes = 24 * (nR + nS) + 8 * n_expected
    pipe_gbs = alg_bytes / dt / 1e9
    # dominant phase: the partition of both relations (read + write both inputs: 16 B/tuple)
    part_bytes = 16 * (nR + nS)
    part_gbs = part_bytes / (ph[0] * 1e-3) / 1e9 if ph[0] > 0 else 0.0
    res = {
        "metric": "join tuples/s & %HBM peak, 2^27 x 2^28 hash join",
        "value": (nR + nS) / dt,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded numpy generator, inputs resident in HBM)",
        "config": {"workload": f"C2 radix hash join |R|=2^{a.log2_r} |S|=2^{a.log2_r + 1} uniform R keys, "
                               f"S zipf theta={a.theta}, selectivity={a.selectivity}",
                   "radix_bits": stats.radix_bits, "passes": stats.passes, "lds_tuples": stats.lds_tuples,
                   "n_out": int(n_expected)},
        "roofline": {"bound": "hbm", "achieved": part_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": part_gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "radix partition (hist+scan+scatter, R and S)"},
        "pipeline": {"alg_bytes": alg_bytes, "achieved_gbs": pipe_gbs, "frac": pipe_gbs / HBM_PEAK_GBS},
    }
    if not a.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(a.cpu_log2_r, a.theta, a.selectivity, a.cpu_threads)
    print(json.dumps(res), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
