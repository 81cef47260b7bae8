"""Benchmarks of the MI355X inner joins.  Prints ONE JSON line.

Default (the headline): C2, a 2^27 x 2^28 radix hash join on one GPU.  One
step = one eo_hash_join over device-resident R and S (inputs already in HBM
when the timed region starts; pairs written to a device buffer).
  value        (|R| + |S|) / seconds per join        [join tuples/s]
  roofline     the dominant kernel (largest device time per join, from HIP
               events recorded around every launch on the join's stream):
               algorithmic bytes per launch / its mean launch time vs 8 TB/s
  pipeline     whole join: 24(|R|+|S|) + 8|out| bytes / join time vs 8 TB/s,
               and each kernel family's algorithmic GB/s
  cpu_baseline the CPU_JOIN hash join restated in oracle/ on a bounded sample
  batched      the metric's second half, from the C5 stream (4096 joins of
               log-uniform sizes): batched throughput, speedup over per-join
               calls, p50/p99 latency under Poisson arrivals (--no-c5 skips it)

Other configurations print their own lines: --config c1 | c3-nlj | c3-hidx |
c4 | c5 | nlj-dist | stream.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


# Algorithmic bytes per launch of each kernel family of the hash join, as a
# function of the tuples the launch processes (n) or of the whole join.
def kernel_alg_bytes(name, nR, nS, n_out):
    per_launch_tuples = (nR + nS) / 2  # each family runs once over R and once over S
    if name in ("part_hist", "seg_hist"):
        return 8 * per_launch_tuples  # read every tuple
    if name == "part_sample":
        return 8 * per_launch_tuples / 16  # one 128-byte line in 16
    if name in ("part_scatter", "seg_scatter", "part_scatter_stable"):
        return 16 * per_launch_tuples  # read + write every tuple
    if name == "probe":
        return 8 * (nR + nS) + 8 * n_out  # read both partitioned relations, write the pairs
    return None


def cpu_baseline_hash(sample_log2_r: int, theta: float, sel: float, threads: int, RS=None):
    """CPU_JOIN hash join (oracle restatement) on a bounded sample of C2, with
    the reference's 128 R tuples per partition (2^20 partitions at |R| = 2^27).
    RS: the GPU run's own inputs when the sample is the full workload."""
    import oracle
    from eating_one_amd import gen

    nR, nS = 1 << sample_log2_r, 1 << (sample_log2_r + 1)
    if RS is not None and len(RS[0]) == nR and len(RS[1]) == nS:
        R, S = RS
    else:
        R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS, theta=theta, selectivity=sel)
    nparts = max(1, nR // 128)
    t0 = time.perf_counter()
    if threads == 1:
        n = oracle.hash_join(R, S, nparts=nparts, count_only=True)
    else:
        n = oracle.hash_join_mt(R, S, nparts=nparts, threads=threads, count_only=True)
    dt = time.perf_counter() - t0
    model, ncpu = host_info()
    # what the timed region produces: the 1-thread restatement counts matches
    # without writing pairs; the pthread one writes every pair into its
    # per-thread buffers (as the reference's threads write results)
    output = ("pairs written to per-thread buffers" if threads > 1
              else "count only: matches counted, no pairs written")
    return {
        "value": (nR + nS) / dt,
        "unit": "tuples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle/cpu_join.c restatement of CPU_JOIN hash_cpu (partition + per-partition nested loop, "
                  f"timer scope hj.cc:209-230{'; pthread version of multi_joins/hash' if threads > 1 else ''}), "
                  f"|R|=2^{sample_log2_r}, |S|=2^{sample_log2_r + 1}, "
                  f"theta={theta}, selectivity={sel}, {nparts} partitions, {n} matches ({output}), {dt:.2f} s on "
                  f"{threads} of {ncpu} host threads ({model})",
    }


def cpu_port(value, unit, sample, dt, threads=1):
    """A cpu_baseline record for an oracle restatement timed on the host."""
    model, ncpu = host_info()
    return {"value": value, "unit": unit, "cores": threads, "kind": "port",
            "sample": f"{sample}, {dt:.2f} s on {threads} of {ncpu} host threads ({model})"}


def cpu_baseline_nlj(S, log2_r_sample: int):
    """CPU_JOIN block_cpu NLJ (timer scope nestLoopJoin.cc:171-173) on the first
    2^log2_r_sample R tuples of C3 against all of S: compares/s."""
    import oracle
    from eating_one_amd import gen

    R, _ = gen.c3()
    r = R[: 1 << log2_r_sample]
    t0 = time.perf_counter()
    n = oracle.nlj_join(r, S, count_only=True)
    dt = time.perf_counter() - t0
    return cpu_port(len(r) * len(S) / dt, "compares/s",
                    f"oracle/cpu_join.c restatement of CPU_JOIN block_cpu, |R| = 2^{log2_r_sample} of C3's R, "
                    f"|S| = {len(S)}, {n} matches (count only: no pairs written)", dt)


def read_traffic(kernel):
    """HBM bytes per launch of a C2 kernel family, from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE runs scaled by the copy-kernel
    calibration recorded in the same file); None if absent."""
    path = os.path.join(ROOT, "profiles", "hash_join_c2_traffic.txt")
    try:
        for line in open(path):
            f = line.split()
            if len(f) >= 6 and f[0] == "traffic" and f[1] == kernel:
                return float(f[5])
    except OSError:
        pass
    return None


def copy_ceiling(ctx):
    g = ctypes.c_double()
    st = ctx.lib.eo_bandwidth_probe(ctx.handle, 1 << 30, 10, ctypes.byref(g))
    return g.value if st == 0 else None


def run_c2(a):
    import numpy as np

    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    nR, nS = 1 << a.log2_r, 1 << (a.log2_r + 1)
    log(f"generating C2 |R|=2^{a.log2_r} |S|=2^{a.log2_r + 1} theta={a.theta} selectivity={a.selectivity}")
    t0 = time.perf_counter()
    R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS, theta=a.theta, selectivity=a.selectivity)
    log(f"generated in {time.perf_counter() - t0:.1f} s")
    ctx = eo.Context(0)
    info = ctx.device_info()
    dR = eo.DeviceArray.from_host(ctx, R)
    dS = eo.DeviceArray.from_host(ctx, S)
    keep = (R, S) if ((not a.no_cpu and a.cpu_log2_r == a.log2_r) or a.host_leg) else None
    del R, S
    kw = dict(radix_bits=a.radix_bits, passes=a.passes, lds_tuples=a.lds, s_tile=a.s_tile)
    n_out, _ = ctx.hash_join_device(dR, nR, dS, nS, None, 0, **kw)
    dout = eo.DeviceArray(ctx, max(n_out, 1) * 8)
    log(f"count: {n_out} pairs")
    stats = _abi.JoinStats()

    def step(st_obj=None):
        n, st = ctx.hash_join_device(dR, nR, dS, nS, dout, n_out, stats=st_obj, **kw)
        if st != _abi.EO_OK or n != n_out:
            raise RuntimeError(f"join failed: status {st}, {n} pairs")

    for _ in range(a.warmup):
        step()
    ctx.sync()
    # timed region: plain joins (no per-launch events)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    ctx.sync()
    dt = (time.perf_counter() - t0) / a.steps
    # kernel breakdown: the same joins again with HIP events around every launch
    kern = {}
    phases = []
    for _ in range(a.steps):
        step(stats)
        for k, (ms, nl) in stats.kernel_table().items():
            tot = kern.setdefault(k, [0.0, 0])
            tot[0] += ms
            tot[1] += nl
        phases.append(stats.ms_total)
    log(f"per join {dt * 1e3:.3f} ms wall (timed, no events); {np.mean(phases):.3f} ms device with per-launch "
        f"events; bits={stats.radix_bits} passes={stats.passes} tiles={stats.tiles}")
    for k, (ms, nl) in sorted(kern.items(), key=lambda x: -x[1][0]):
        log(f"  {k:20s} {ms / a.steps:8.3f} ms/join  {nl / a.steps:.0f} launches/join  {ms / nl:8.3f} ms/launch")
    # the deterministic join (ordered partition, count -> scan -> write in a fixed
    # table order): the same inputs, reported beside the fused join, not as `value`
    det_kw = dict(kw, deterministic=True)
    for it in range(2 + a.steps):
        if it == 2:
            ctx.sync()
            t0 = time.perf_counter()
        n, st = ctx.hash_join_device(dR, nR, dS, nS, dout, n_out, **det_kw)
        if st != _abi.EO_OK or n != n_out:
            raise RuntimeError(f"deterministic join failed: status {st}, {n} pairs")
    ctx.sync()
    det_dt = (time.perf_counter() - t0) / a.steps
    log(f"deterministic join: {det_dt * 1e3:.3f} ms")
    # the configuration's other partition plans (one pass at 8 and 11 bits, two
    # passes at 12 and 13), a few joins each, beside the default plan's `value`
    plans = []
    if not a.no_plans and a.log2_r == 27:
        for bits, passes in ((8, 1), (11, 1), (12, 2), (13, 2)):
            pkw = dict(kw, radix_bits=bits, passes=passes)
            pst = _abi.JoinStats()
            for it in range(4):
                if it == 1:
                    ctx.sync()
                    t0 = time.perf_counter()
                n, st = ctx.hash_join_device(dR, nR, dS, nS, dout, n_out, stats=pst if it == 0 else None, **pkw)
                if st != _abi.EO_OK or n != n_out:
                    raise RuntimeError(f"{bits}-bit {passes}-pass join failed: status {st}, {n} pairs")
            ctx.sync()
            pdt = (time.perf_counter() - t0) / 3
            plans.append({"radix_bits": bits, "passes": passes, "ms_per_join": pdt * 1e3,
                          "tuples_per_s": (nR + nS) / pdt, "sampled": int(pst.sampled)})
            log(f"plan {bits} bits, {passes} pass(es): {pdt * 1e3:.3f} ms")
    dom = max(kern.items(), key=lambda x: x[1][0])
    dom_name, (dom_ms, dom_launches) = dom
    dom_mean_ms = dom_ms / dom_launches
    dom_bytes = kernel_alg_bytes(dom_name, nR, nS, n_out)
    achieved = dom_bytes / (dom_mean_ms * 1e-3) / 1e9 if dom_bytes else None
    traffic = read_traffic(dom_name) if a.log2_r == 27 else None
    alg_bytes = 24 * (nR + nS) + 8 * n_out
    pipe = alg_bytes / dt / 1e9
    # HBM bytes a join actually moves: the committed FETCH_SIZE / WRITE_SIZE
    # figures per launch of each kernel family x this run's launches per join
    # (the families carry all but a few MB of a join's traffic)
    join_hbm = None
    if a.log2_r == 27:
        per = {k: read_traffic(k) for k in kern}
        if per and all(per[k] is not None for k in kern):
            join_hbm = sum(per[k] * kern[k][1] / a.steps for k in kern)
    ceiling = copy_ceiling(ctx)
    res = {
        "metric": "join tuples/s & %HBM peak, 2^27\u22c82^28 hash join; p99 latency, batched joins",
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
        "data": "synthetic: seeded numpy generator (R = permutation of [0,|R|), S keys drawn from R keys), "
                "inputs resident in HBM",
        "config": {"workload": f"C2 radix hash join |R|=2^{a.log2_r} |S|=2^{a.log2_r + 1}, theta={a.theta}, "
                               f"selectivity={a.selectivity}",
                   "radix_bits": int(stats.radix_bits), "passes": int(stats.passes),
                   "partitions": ("exact (histogram pass)", "sized from a 1/16 sample",
                                  "sample too small, redone exactly")[min(int(stats.sampled), 2)],
                   "n_out": int(n_out),
                   "device": info["arch"], "cus": info["cus"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "kernel": dom_name, "alg_bytes_per_launch": dom_bytes, "ms_per_launch": dom_mean_ms},
        "pipeline": {"alg_bytes": alg_bytes, "achieved_gbs": pipe, "frac": pipe / HBM_PEAK_GBS,
                     "hbm_bytes_per_join": join_hbm,
                     "hbm_gbs": join_hbm / dt / 1e9 if join_hbm else None,
                     "hbm_frac": join_hbm / dt / 1e9 / HBM_PEAK_GBS if join_hbm else None,
                     "kernels_ms_per_join": {k: v[0] / a.steps for k, v in kern.items()},
                     # every kernel family: algorithmic bytes per launch / mean launch time
                     "kernels_gbs": {k: kernel_alg_bytes(k, nR, nS, n_out) / (v[0] / v[1] * 1e-3) / 1e9
                                     for k, v in kern.items() if kernel_alg_bytes(k, nR, nS, n_out)}},
        "copy_ceiling_gbs": ceiling,
        "deterministic": {"ms_per_join": det_dt * 1e3, "tuples_per_s": (nR + nS) / det_dt,
                          "note": "same rows in the same order on every run; not the headline value"},
    }
    if plans:
        res["plans"] = plans
    if a.host_leg:
        # the PCIe-inclusive rate: host R and S in, pairs back to host memory (not `value`)
        hR, hS = keep
        hout = np.empty((n_out, 2), dtype=np.int32)
        nn = ctypes.c_uint64()
        oh = ctx.opts(where=_abi.EO_HOST, **kw)
        tl = []
        for it in range(3):
            t1 = time.perf_counter()
            st = ctx.lib.eo_hash_join(ctx.handle, hR.ctypes.data, nR, hS.ctypes.data, nS, hout.ctypes.data, n_out,
                                      ctypes.byref(nn), ctypes.byref(oh))
            tl.append(time.perf_counter() - t1)
            assert st == 0 and nn.value == n_out
        res["pcie_inclusive"] = {"ms_per_join": min(tl) * 1e3, "tuples_per_s": (nR + nS) / min(tl),
                                 "note": "host R, S in and pairs out through pageable host memory"}
        log(f"host path (PCIe-inclusive): {min(tl) * 1e3:.1f} ms per join")
    ctx.close()
    if not a.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline_hash(a.cpu_log2_r, a.theta, a.selectivity, a.cpu_threads, keep)
        if a.cpu_mt_threads > 1:
            # the multi_joins-style pthread restatement on the box's host share, reported beside it
            mt = cpu_baseline_hash(a.cpu_log2_r, a.theta, a.selectivity, a.cpu_mt_threads, keep)
            res["cpu_baseline"]["multi_thread"] = {k: mt[k] for k in ("value", "unit", "cores", "sample")}
    return res


def run_stream(a):
    """C2 sizes held in host memory, joined by eo_stream_join (R resident, S streamed
    in chunks, rows copied back): the PCIe-bound out-of-core path."""
    import numpy as np

    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    nR, nS = 1 << a.log2_r, 1 << (a.log2_r + 1)
    log(f"generating C2 |R|=2^{a.log2_r} |S|=2^{a.log2_r + 1} in host memory")
    R, S = gen.c2(seed=0xE01C2, nR=nR, nS=nS, theta=a.theta, selectivity=a.selectivity)
    ctx = eo.Context(0)
    n_out = ctx.stream_join(R, S, s_chunk=a.s_chunk, count_only=True)
    hR, hS, hO = eo.HostArray(ctx, R.nbytes), eo.HostArray(ctx, S.nbytes), eo.HostArray(ctx, 8 * max(n_out, 1))
    vR, vS, vO = hR.view(np.int32, R.shape), hS.view(np.int32, S.shape), hO.view(np.int32, (max(n_out, 1), 2))
    vR[:] = R
    vS[:] = S
    # PCIe rates of plain pinned copies (1 GiB each way) for the roofline
    cb = min(1 << 30, S.nbytes, 8 * max(n_out, 1))
    dbuf = eo.DeviceArray(ctx, cb)
    rates = {}
    for name, fn, dst, src in (("h2d", ctx.lib.eo_copy_h2d, dbuf.ptr, hS.ptr), ("d2h", ctx.lib.eo_copy_d2h, hO.ptr, dbuf.ptr)):
        fn(ctx.handle, dst, src, cb)
        t0 = time.perf_counter()
        for _ in range(3):
            assert fn(ctx.handle, dst, src, cb) == 0
        rates[name] = 3 * cb / (time.perf_counter() - t0) / 1e9
    dbuf.close()
    legs = {}
    for leg, (r, s_, o) in (("pinned", (vR, vS, vO)), ("pageable", (R, S, np.empty((max(n_out, 1), 2), np.int32)))):
        times = []
        st = _abi.StreamStats()
        for it in range(a.warmup + a.steps):
            t0 = time.perf_counter()
            n, status = ctx.stream_join(r, s_, s_chunk=a.s_chunk, out=o, stats=st)
            dt = time.perf_counter() - t0
            assert status == _abi.EO_OK and n == n_out
            if it >= a.warmup:
                times.append(dt)
        ms = 1e3 * sum(times) / len(times)
        legs[leg] = {"ms": ms, "tuples_per_s": (nR + nS) / (ms * 1e-3),
                     "pcie_gbs": (st.h2d_bytes + st.d2h_bytes) / (ms * 1e-3) / 1e9,
                     "h2d_bytes": st.h2d_bytes, "d2h_bytes": st.d2h_bytes, "s_chunks": st.s_chunks,
                     "pinned_flags": st.pinned}
        log(f"{leg}: {ms:.1f} ms per join, {legs[leg]['pcie_gbs']:.1f} GB/s over PCIe")
    # eo_hash_join on the same pageable host buffers (it streams S itself at this size)
    o = ctx.opts()
    out = np.empty((max(n_out, 1), 2), np.int32)
    n = ctypes.c_uint64()
    times = []
    for it in range(1 + min(a.steps, 3)):
        t0 = time.perf_counter()
        assert ctx.lib.eo_hash_join(ctx.handle, R.ctypes.data, nR, S.ctypes.data, nS, out.ctypes.data, n_out,
                                    ctypes.byref(n), ctypes.byref(o)) == 0
        if it:
            times.append(time.perf_counter() - t0)
    host_join_ms = 1e3 * sum(times) / len(times)
    pin = legs["pinned"]
    moved = pin["h2d_bytes"] + pin["d2h_bytes"]
    # the bound: each direction at its plain-copy rate, both directions at once
    floor_ms = 1e3 * max(pin["h2d_bytes"] / (rates["h2d"] * 1e9), pin["d2h_bytes"] / (rates["d2h"] * 1e9))
    res = {"metric": "streamed join tuples/s, C2 2^27 x 2^28 in host memory", "value": pin["tuples_per_s"],
           "unit": "tuples/s", "ms_per_step": pin["ms"], "steps": a.steps, "warmup": a.warmup,
           "config": {"workload": f"C2 hash join |R|=2^{a.log2_r} |S|=2^{a.log2_r + 1} from host memory, "
                                  f"s_chunk={a.s_chunk or 1 << 24}", "n_out": n_out},
           "legs": legs, "eo_hash_join_host_ms": host_join_ms, "pcie_copy_gbs": rates,
           "roofline": {"bound": "pcie", "achieved": moved / (pin["ms"] * 1e-3) / 1e9,
                        "peak": rates["h2d"] + rates["d2h"], "unit": "GB/s",
                        "frac_of_duplex_floor": floor_ms / pin["ms"]}}
    hO.close()
    hS.close()
    hR.close()
    ctx.close()
    return res


def run_c3_nlj(a):
    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    R, S = gen.c3()
    ctx = eo.Context(0)
    info = ctx.device_info()
    dR, dS = eo.DeviceArray.from_host(ctx, R), eo.DeviceArray.from_host(ctx, S)
    L = ctx.lib
    n = ctypes.c_uint64()
    o = ctx.opts(where=_abi.EO_DEVICE)
    assert L.eo_nlj_join(ctx.handle, dR.ptr, len(R), dS.ptr, len(S), None, 0, ctypes.byref(n), ctypes.byref(o)) == 0
    dout = eo.DeviceArray(ctx, n.value * 8)
    stats = _abi.JoinStats()
    o = ctx.opts(where=_abi.EO_DEVICE, stats=stats)
    kern = {}
    for it in range(a.warmup + a.steps):
        st = L.eo_nlj_join(ctx.handle, dR.ptr, len(R), dS.ptr, len(S), dout.ptr, n.value, ctypes.byref(n),
                           ctypes.byref(o))
        assert st == 0
        if it >= a.warmup:
            for k, (ms, nl) in stats.kernel_table().items():
                kern[k] = kern.get(k, 0.0) + ms
    ms = {k: v / a.steps for k, v in kern.items()}
    # the deterministic count -> scan -> write variant, for comparison
    od = ctx.opts(where=_abi.EO_DEVICE, stats=stats, deterministic=True)
    det = []
    for it in range(a.warmup + a.steps):
        assert L.eo_nlj_join(ctx.handle, dR.ptr, len(R), dS.ptr, len(S), dout.ptr, n.value, ctypes.byref(n),
                             ctypes.byref(od)) == 0
        if it >= a.warmup:
            det.append(stats.ms_total)
    compares = len(R) * len(S)
    peak = info["cus"] * 64 * info["clock_khz"] * 1e3  # one 32-bit compare per lane per clock
    dom = max(ms.items(), key=lambda x: x[1])
    total_ms = sum(ms.values())
    rate = compares / (dom[1] * 1e-3)
    # measured issue rates: the NLJ's own compare loop without hits, and plain
    # integer adds (eo_valu_probe), best over workgroups per CU
    measured = {}
    for name, kind in (("compare_loop", _abi.EO_VALU_COMPARE), ("int_add", _abi.EO_VALU_ADD)):
        best = 0.0
        for bpc in (2, 4, 8):
            r = ctypes.c_double()
            assert L.eo_valu_probe(ctx.handle, kind, bpc, 4 if kind == _abi.EO_VALU_COMPARE else 64, 3,
                                   ctypes.byref(r)) == 0
            best = max(best, r.value)
        measured[name] = best
        log(f"measured {name}: {best / 1e12:.2f} T/s ({best / peak:.2f} of 1 per lane per clock)")
    res = {"metric": "NLJ compares/s, 2^18 x 2^18 (C3)", "value": compares / (total_ms * 1e-3),
           "unit": "compares/s", "ms_per_step": total_ms, "n_out": n.value, "kernels_ms": ms,
           "deterministic_ms": sum(det) / len(det),
           "roofline": {"bound": "valu", "achieved": rate / 1e9, "peak": peak / 1e9, "unit": "Gcompares/s",
                        "frac": rate / peak, "kernel": dom[0],
                        "measured_compare_loop": measured["compare_loop"] / 1e9,
                        "measured_int_add": measured["int_add"] / 1e9,
                        "frac_of_compare_loop": rate / measured["compare_loop"]}}
    ctx.close()
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline_nlj(S, 16)
    return res


def run_nlj_dist(a):
    """Distance-predicate NLJ (the reference's CPU_JOIN/3d_cpu predicate) on
    2^17 x 2^17 random 3-D points; FP64 roofline: 8 flops per compared pair
    (3 subtractions, 3 products, 2 sums) against MI355X's 78.6 TFLOP/s."""
    import numpy as np

    import eating_one_amd as eo
    from eating_one_amd import _abi

    n = 1 << 17
    rng = np.random.default_rng(0xE01D3)
    pt = eo.Context.POINT3
    R, S = np.zeros(n, dtype=pt), np.zeros(n, dtype=pt)
    R["id"], S["id"] = np.arange(n), np.arange(n)
    R["v"] = rng.uniform(0, 64, size=(n, 3))
    S["v"] = rng.uniform(0, 64, size=(n, 3))
    ctx = eo.Context(0)
    dR, dS = eo.DeviceArray.from_host(ctx, R), eo.DeviceArray.from_host(ctx, S)
    L = ctx.lib
    cnt = ctypes.c_uint64()
    o = ctx.opts(where=_abi.EO_DEVICE)
    assert L.eo_nlj_dist_join(ctx.handle, dR.ptr, n, dS.ptr, n, 1.0, None, 0, ctypes.byref(cnt), ctypes.byref(o)) == 0
    dout = eo.DeviceArray(ctx, max(cnt.value, 1) * 8)
    stats = _abi.JoinStats()
    o = ctx.opts(where=_abi.EO_DEVICE, stats=stats)
    t = []
    for it in range(a.warmup + a.steps):
        assert L.eo_nlj_dist_join(ctx.handle, dR.ptr, n, dS.ptr, n, 1.0, dout.ptr, cnt.value, ctypes.byref(cnt),
                                  ctypes.byref(o)) == 0
        if it >= a.warmup:
            t.append(stats.ms_total)
    ms = float(np.median(t))
    pairs = float(n) * n
    tflops = 8 * pairs / (ms * 1e-3) / 1e12
    res = {"metric": "distance NLJ compared pairs/s, 2^17 x 2^17 3-D points", "value": pairs / (ms * 1e-3),
           "unit": "pairs/s", "ms_per_step": ms, "n_out": cnt.value,
           "roofline": {"bound": "fp64", "achieved": tflops, "peak": 78.6, "unit": "TFLOP/s", "frac": tflops / 78.6,
                        "kernel": "nlj_dist"}}
    ctx.close()
    if not a.no_cpu:
        import oracle
        r = R[: 1 << 14]
        t0 = time.perf_counter()
        m = oracle.nlj_dist_join(r, S, 1.0, count_only=True)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = cpu_port(len(r) * float(n) / dt, "pairs/s",
                                       f"oracle/cpu_join.c restatement of CPU_JOIN 3d_cpu, |R| = 2^14 of the "
                                       f"2^17 points, |S| = 2^17, {m} matches (count only: no pairs written)", dt)
    return res


def run_c3_hidx(a):
    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    R, S = gen.c3()
    ctx = eo.Context(0)
    stats = _abi.JoinStats()
    builds = []
    for it in range(1 + 5):  # the first build also loads the kernels' code: not timed
        ix = ctx.hidx_build(R, stats=stats)
        if it:
            builds.append(stats.ms_build)
        if it < 5:
            ix.free()
    build_ms = sorted(builds)[len(builds) // 2]
    dS = eo.DeviceArray.from_host(ctx, S)
    L = ctx.lib
    n = ctypes.c_uint64()
    o = ctx.opts(where=_abi.EO_DEVICE)
    assert L.eo_index_join(ctx.handle, ix.handle, dS.ptr, len(S), None, 0, ctypes.byref(n), ctypes.byref(o)) == 0
    dout = eo.DeviceArray(ctx, n.value * 8)
    o = ctx.opts(where=_abi.EO_DEVICE, stats=stats)
    t = []
    for it in range(a.warmup + a.steps):
        assert L.eo_index_join(ctx.handle, ix.handle, dS.ptr, len(S), dout.ptr, n.value, ctypes.byref(n),
                               ctypes.byref(o)) == 0
        if it >= a.warmup:
            t.append(stats.ms_total)
    ms = sum(t) / len(t)
    alg = 8 * len(R) + 16 * len(R) + 8 * len(S) + 8 * n.value
    res = {"metric": "hash-index join probes/s, 2^18 x 2^18 (C3)", "value": len(S) / (ms * 1e-3),
           "unit": "probes/s", "ms_per_step": ms, "build_ms": build_ms, "n_out": n.value,
           "pipeline": {"alg_bytes": alg, "achieved_gbs": alg / (ms + build_ms) / 1e6}}
    ix.free()
    ctx.close()
    if not a.no_cpu:
        import oracle
        t0 = time.perf_counter()
        m = oracle.hidx_join(R, S, count_only=True)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = cpu_port(len(S) / dt, "probes/s",
                                       f"oracle/cpu_join.c restatement of CPU_JOIN inner_cpu (2^27-bucket index "
                                       f"build + probe, timer scope inner-idx-bnlj.cc:142-186), full C3, {m} matches "
                                       f"(count only: no pairs written)",
                                       dt)
    return res


def run_c4(a):
    import numpy as np

    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    nR, nS = 1 << a.c4_log2_r, 1 << (a.c4_log2_r + 1)
    log(f"generating C4 |R|=2^{a.c4_log2_r} lognormal(0,2) |S|=2^{a.c4_log2_r + 1}")
    R, S = gen.c4(nR=nR, nS=nS)
    ctx = eo.Context(0)
    stats = _abi.JoinStats()
    builds = []
    for it in range(1 + 3):  # the first build also loads the kernels' code: not timed
        ix = ctx.sidx_build(R, learned=True, stats=stats)
        if it:
            builds.append(stats.ms_build)
        if it < 3:
            ix.free()
    info = ix.info()
    info["build_ms"] = sorted(builds)[len(builds) // 2]
    dS = eo.DeviceArray.from_host(ctx, S)
    L = ctx.lib
    pos = eo.DeviceArray(ctx, nS * 8)
    res = {"metric": "sorted-index probes/s, learned vs binary search (C4)", "unit": "probes/s",
           "index": info, "n": {"R": nR, "S": nS}}
    for method, name in ((_abi.EO_LOOKUP_BINARY, "binary"), (_abi.EO_LOOKUP_LEARNED, "learned")):
        o = ctx.opts(where=_abi.EO_DEVICE, stats=stats)
        t = []
        for it in range(a.warmup + a.steps):
            assert L.eo_index_lookup(ctx.handle, ix.handle, dS.ptr, nS, method, pos.ptr, ctypes.byref(o)) == 0
            if it >= a.warmup:
                t.append(stats.ms_total)
        res[f"{name}_lookup_ms"] = float(np.median(t))
        n = ctypes.c_uint64()
        flags = _abi.EO_F_BINARY_SEARCH if method == _abi.EO_LOOKUP_BINARY else 0
        o = ctx.opts(where=_abi.EO_DEVICE, flags=flags)
        assert L.eo_index_join(ctx.handle, ix.handle, dS.ptr, nS, None, 0, ctypes.byref(n), ctypes.byref(o)) == 0
        dout = eo.DeviceArray(ctx, max(n.value, 1) * 8)
        o = ctx.opts(where=_abi.EO_DEVICE, flags=flags, stats=stats)
        t = []
        for it in range(a.warmup + a.steps):
            assert L.eo_index_join(ctx.handle, ix.handle, dS.ptr, nS, dout.ptr, n.value, ctypes.byref(n),
                                   ctypes.byref(o)) == 0
            if it >= a.warmup:
                t.append(stats.ms_total)
        res[f"{name}_join_ms"] = float(np.median(t))
        res["n_out"] = n.value
        dout.close()
    res["value"] = nS / (res["learned_lookup_ms"] * 1e-3)
    res["speedup_lookup"] = res["binary_lookup_ms"] / res["learned_lookup_ms"]
    res["speedup_join"] = res["binary_join_ms"] / res["learned_join_ms"]
    ix.free()
    ctx.close()
    if not a.no_cpu:
        import oracle
        r, s_ = gen.c4(nR=1 << 24, nS=1 << 25)  # 1/4 of the workload
        t0 = time.perf_counter()
        oracle.sidx_build(r)
        build_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        m = oracle.sidx_join(r, s_, count_only=True)
        dt = time.perf_counter() - t0
        probe_s = max(dt - build_s, 1e-9)
        res["cpu_baseline"] = cpu_port(len(s_) / probe_s, "probes/s",
                                       f"oracle/cpu_join.c restatement of CPU_JOIN sortIndex_cpu (quicksort build "
                                       f"{build_s:.2f} s, then binary-search probes; timer scopes "
                                       f"inner-idx-bnlj.cc:204-238), C4 shape at |R| = 2^24, |S| = 2^25, "
                                       f"{m} matches (count only: no pairs written)", dt)
        if a.cpu_mt_threads > 1:
            t0 = time.perf_counter()
            m2 = oracle.sidx_join_mt(r, s_, threads=a.cpu_mt_threads, count_only=True)
            dt2 = time.perf_counter() - t0
            mt = cpu_port(len(s_) / max(dt2 - build_s, 1e-9), "probes/s",
                          f"pthread restatement of multi_joins/sortindex (serial quicksort build, then "
                          f"{a.cpu_mt_threads} probe threads), same sample, {m2} matches (count only)", dt2,
                          a.cpu_mt_threads)
            res["cpu_baseline"]["multi_thread"] = {k: mt[k] for k in ("value", "unit", "cores", "sample")}
    return res


def run_c1(a):
    """C1: the CPU_JOIN path's own shape (2^16 x 2^20, every S tuple matches
    once), launch-bound on the GPU; the oracle on the full inputs beside it."""
    import numpy as np

    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    R, S = gen.c1()
    ctx = eo.Context(0)
    dR, dS = eo.DeviceArray.from_host(ctx, R), eo.DeviceArray.from_host(ctx, S)
    n, _ = ctx.hash_join_device(dR, len(R), dS, len(S), None, 0)
    dout = eo.DeviceArray(ctx, 8 * n)
    t = []
    for it in range(a.warmup + a.steps):
        t0 = time.perf_counter()
        n2, st = ctx.hash_join_device(dR, len(R), dS, len(S), dout, n)
        if it >= a.warmup:
            t.append(time.perf_counter() - t0)
        assert st == _abi.EO_OK and n2 == n
    ms = 1e3 * float(np.median(t))
    alg = 24 * (len(R) + len(S)) + 8 * n
    res = {"metric": "join tuples/s, C1 2^16 x 2^20", "value": (len(R) + len(S)) / (ms * 1e-3),
           "unit": "tuples/s", "ms_per_step": ms, "n_out": n, "steps": a.steps, "warmup": a.warmup,
           "pipeline": {"alg_bytes": alg, "achieved_gbs": alg / (ms * 1e-3) / 1e9}}
    for d in (dout, dS, dR):
        d.close()
    ctx.close()
    if not a.no_cpu:
        import oracle
        t0 = time.perf_counter()
        m = oracle.hash_join(R, S, count_only=True)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = cpu_port((len(R) + len(S)) / dt, "tuples/s",
                                       f"oracle/cpu_join.c restatement of CPU_JOIN hash_cpu (2^20 partitions, "
                                       f"timer scope hj.cc:209-230), full C1, {m} matches (count only: no pairs written)", dt)
    return res


def run_c5(a):
    """C5: a stream of joins of log-uniform sizes, served three ways on the same
    device-resident requests: per-join eo_hash_join calls (closed loop), batched
    (closed loop), and batched under Poisson arrivals (latency)."""
    import numpy as np

    import eating_one_amd as eo
    from eating_one_amd import _abi, gen

    rng = np.random.default_rng((a.seed if a.seed is not None else 0xE01C5) + 1)  # arrival times
    n_req = a.c5_requests
    ctx = eo.Context(0, hbm_budget=int(a.c5_budget))
    t0 = time.time()
    # 64 resident build tables; 25 % of the requests reuse one of them
    stream = gen.C5Stream(ctx, n_req=n_req, lo=a.c5_lo, hi=a.c5_hi, n_tables=a.c5_tables, reuse=a.c5_reuse,
                          seed=a.seed if a.seed is not None else 0xE01C5)
    reqs, nR, nS = stream.reqs, stream.nR, stream.nS
    log(f"C5: {n_req} requests, |R|,|S| = 2^U, U~U[{a.c5_lo},{a.c5_hi}]; R {8 * stream.r_off[-1] / 1e9:.1f} GB, "
        f"S {8 * stream.s_off[-1] / 1e9:.1f} GB, out {8 * stream.s_off[-1] / 1e9:.1f} GB; "
        f"{int(stream.use_table.sum())} on tables")
    log(f"generated in {time.time() - t0:.1f} s")
    tuples = int(nR.sum() + nS.sum())
    kw = dict(max_batch=a.c5_max_batch, slots=a.c5_slots)
    out = {}
    counts = None
    modes = [m for m in (("per_join", _abi.EO_SERVE_PER_JOIN), ("batched", _abi.EO_SERVE_BATCHED))
             if a.c5_modes in ("both", m[0])]
    for name, mode in modes:
        # warm-up until two runs agree (at most four): the batch slots' arenas
        # reach their steady size over the first runs, and on device memory
        # another process has just freed each of those growths costs far more
        # than on fresh memory (measured: 1.5 and 1.2 s, then 75 ms)
        prev = None
        for _ in range(4):
            _, _, wst = ctx.serve(reqs, None, mode=mode, **kw)
            if prev is not None and abs(wst["wall_ms"] - prev) <= 0.15 * prev:
                break
            prev = wst["wall_ms"]
        walls = []
        for _ in range(max(1, a.c5_reps)):
            res, lat, st = ctx.serve(reqs, None, mode=mode, **kw)
            walls.append(st["wall_ms"])
        got = [r[1] for r in res]
        assert all(r[2] == 0 for r in res), name
        if counts is None:
            counts = got
        assert got == counts, "per-join and batched counts differ"
        out[name] = {"wall_ms": float(np.median(walls)), "tuples_per_s": tuples / (np.median(walls) * 1e-3),
                     "batches": st["batches"], "largest_batch": st["largest_batch"]}
        log(f"{name}: {np.median(walls):.1f} ms for {n_req} requests ({tuples / np.median(walls) / 1e6:.2f} G tuples/s)")
    if a.c5_modes != "both":  # one mode only (profiling): closed loop only
        stream.free()
        ctx.close()
        return {"metric": "C5 join stream, one mode", "closed_loop": out}
    alg_bytes = 24 * tuples + 8 * int(sum(counts))
    capacity_qps = n_req / (out["batched"]["wall_ms"] * 1e-3)
    qps = a.c5_qps if a.c5_qps else round(0.7 * capacity_qps, -2)
    arrival = np.cumsum(rng.exponential(1000.0 / qps, size=n_req))
    lat_res = {}
    for name, mode in (("batched", _abi.EO_SERVE_BATCHED), ("per_join", _abi.EO_SERVE_PER_JOIN)):
        res, lat, st = ctx.serve(reqs, arrival, mode=mode, **kw)
        assert [r[1] for r in res] == counts
        lat_res[name] = {"p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
                         "mean_ms": st["mean_ms"], "wall_ms": st["wall_ms"], "batches": st["batches"],
                         "peak_admitted_gb": st["peak_admitted"] / 1e9}
        log(f"{name} at {qps:.0f} req/s: p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms")
    res = {"metric": "C5 join stream throughput, batched persistent kernel", "value": out["batched"]["tuples_per_s"],
           "unit": "join tuples/s", "higher_is_better": True, "dtype": "int32", "data": "synthetic (device generator)",
           "config": {"workload": f"C5 stream: {n_req} joins, |R|,|S|=2^U U~U[{a.c5_lo},{a.c5_hi}], uniform keys, "
                                  f"selectivity U[0.1,1], {a.c5_tables} resident tables reused by "
                                  f"{a.c5_reuse:.0%} of requests, HBM budget {a.c5_budget / 1e9:.0f} GB",
                      "requests": n_req, "tuples": tuples, "n_out": int(sum(counts)), "max_batch": a.c5_max_batch,
                      "slots": a.c5_slots},
           "closed_loop": out, "speedup_batched_vs_per_join": out["per_join"]["wall_ms"] / out["batched"]["wall_ms"],
           # the C2 formula per request (read, write and re-read both inputs, write the pairs)
           "roofline": {"bound": "hbm", "alg_bytes": alg_bytes,
                        "achieved": alg_bytes / (out["batched"]["wall_ms"] * 1e-3) / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": alg_bytes / (out["batched"]["wall_ms"] * 1e-3) / 8e12},
           "poisson": {"qps": qps, **{k: v for k, v in lat_res.items()}}}
    stream.free()
    res["graph"] = graph_leg(ctx, a)
    ctx.close()
    return res


def graph_leg(ctx, a, n_req=1024, lo=10, hi=16, reps=20):
    """A fixed batch of small joins (C5 shapes, |R|, |S| = 2^U, U ~ U[lo, hi])
    run eagerly (eo_batch_submit: host planning, descriptor upload, launches)
    and replayed from its hipGraph (eo_batch_replay: one graph launch), each
    followed by eo_batch_wait; medians of `reps` runs, rows checked equal."""
    import numpy as np

    from eating_one_amd import gen

    s = gen.C5Stream(ctx, n_req=n_req, lo=lo, hi=hi, n_tables=16, reuse=0.25, seed=0x6EA9)
    be, bg = ctx.batch(), ctx.batch()  # a new submit discards a batch object's graph: one object each
    eager, graph = [], []
    first = be.run(s.reqs)
    assert bg.run(s.reqs) == first
    bg.replay()
    assert bg.wait() == first  # capture, instantiate and the first replay
    for _ in range(reps):
        t0 = time.perf_counter()
        res = be.run(s.reqs)
        eager.append((time.perf_counter() - t0) * 1e3)
        assert res == first
        t0 = time.perf_counter()
        bg.replay()
        res = bg.wait()
        graph.append((time.perf_counter() - t0) * 1e3)
        assert res == first
    be.free()
    bg.free()
    tuples = int(s.nR.sum() + s.nS.sum())
    s.free()
    e, g = float(np.median(eager)), float(np.median(graph))
    log(f"graph leg: {n_req} joins of 2^{lo}..2^{hi}: eager {e:.3f} ms, graph replay {g:.3f} ms")
    return {"workload": f"one batch of {n_req} joins, |R|,|S|=2^U U~U[{lo},{hi}], 16 resident tables for 25%",
            "tuples": tuples, "eager_ms": e, "graph_replay_ms": g, "speedup": e / g}


def run_copy(a):
    """The HBM ceiling the join kernels are compared against: a sweep of the
    bandwidth probe (copy, read-only and write-only passes over 1 and 4 GiB;
    16-byte words in flight per thread, workgroup size, workgroups per CU,
    streaming or normal loads and stores).  Prints every point to stderr and
    one JSON line with the best of each mode."""
    import eating_one_amd as eo
    from eating_one_amd import _abi

    ctx = eo.Context(0)
    best = {}
    rows = []
    modes = {"copy": _abi.EO_PROBE_COPY, "read": _abi.EO_PROBE_READ, "write": _abi.EO_PROBE_WRITE}
    for size in (1 << 30, 1 << 32):
        for mname, mode in modes.items():
            for unroll in (1, 2, 4, 8):
                for threads in (256, 512, 1024):
                    for bpc in (0, 4, 8, 16, 32):
                        if bpc and bpc * threads > 2048 * 4:  # more threads than a CU holds
                            continue
                        for flags in (0, 1, 2, 3):
                            if mode == _abi.EO_PROBE_READ and flags & 2 or mode == _abi.EO_PROBE_WRITE and flags & 1:
                                continue
                            if bpc == 0 and unroll > 1:  # one word per thread: nothing to unroll
                                continue
                            o = _abi.ProbeOpts(ctypes.sizeof(_abi.ProbeOpts), mode, size, unroll, threads, bpc,
                                               flags, 5, 0)
                            g = ctypes.c_double()
                            st = ctx.lib.eo_bandwidth_probe_ex(ctx.handle, ctypes.byref(o), ctypes.byref(g))
                            if st != 0:
                                raise RuntimeError(f"eo_bandwidth_probe_ex failed ({st})")
                            key = (mname, size)
                            row = {"mode": mname, "gib": size >> 30, "unroll": unroll, "threads": threads,
                                   "blocks_per_cu": bpc, "nt_load": bool(flags & 1), "nt_store": bool(flags & 2),
                                   "gbs": round(g.value, 1)}
                            rows.append(row)
                            if key not in best or g.value > best[key]["gbs"]:
                                best[key] = row
            log(f"{mname} {size >> 30} GiB best: {best[(mname, size)]}")
    # the radix scatter's write pattern: sequential reads, writes in runs of
    # run_bytes sent to scattered places (1 GiB)
    scatter = {}
    for run in (128, 256, 512, 1024, 2048, 4096, 16384, 65536):
        top = None
        for unroll, threads, bpc in ((1, 256, 0), (4, 256, 8), (4, 1024, 2), (2, 512, 8)):
            for flags in (1, 3):
                o = _abi.ProbeOpts(ctypes.sizeof(_abi.ProbeOpts), _abi.EO_PROBE_SCATTER, 1 << 30, unroll, threads,
                                   bpc, flags, 5, run)
                g = ctypes.c_double()
                if ctx.lib.eo_bandwidth_probe_ex(ctx.handle, ctypes.byref(o), ctypes.byref(g)) != 0:
                    raise RuntimeError("eo_bandwidth_probe_ex (scatter) failed")
                row = {"mode": "scatter", "run_bytes": run, "unroll": unroll, "threads": threads,
                       "blocks_per_cu": bpc, "nt_store": bool(flags & 2), "gbs": round(g.value, 1)}
                rows.append(row)
                if top is None or g.value > top["gbs"]:
                    top = row
        scatter[run] = top
        log(f"scatter runs of {run} B best: {top}")
    ctx.close()
    top = {f"{m}_{sz >> 30}gib": r for (m, sz), r in best.items()}
    return {"metric": "HBM bandwidth ceiling (bandwidth probe sweep)", "unit": "GB/s",
            "value": max(r["gbs"] for (m, _), r in best.items() if m == "copy"), "best": top,
            "scatter_by_run_bytes": scatter, "points": len(rows),
            "sweep": rows if a.copy_all else None}


def other_configs(a):
    """The other configurations' GPU numbers, measured in the same run as the
    headline (their CPU baselines and full records: bench.py --config ...)."""
    b = argparse.Namespace(**vars(a))
    b.no_cpu = True
    b.steps, b.warmup = 5, 2
    out = {}
    c1 = run_c1(b)
    out["c1"] = {"workload": "C1 hash join 2^16 x 2^20", "ms_per_join": c1["ms_per_step"], "tuples_per_s": c1["value"]}
    nlj = run_c3_nlj(b)
    out["c3_nlj"] = {"workload": "C3 block NLJ 2^18 x 2^18", "ms_per_join": nlj["ms_per_step"],
                     "compares_per_s": nlj["value"], "frac_of_nominal": nlj["roofline"]["frac"],
                     "frac_of_compare_loop": nlj["roofline"].get("frac_of_compare_loop")}
    hidx = run_c3_hidx(b)
    out["c3_hidx"] = {"workload": "C3 hash-index join 2^18 x 2^18", "ms_per_probe_pass": hidx["ms_per_step"],
                      "probes_per_s": hidx["value"], "build_ms": hidx.get("build_ms")}
    c4 = run_c4(b)
    out["c4"] = {"workload": "C4 sorted index 2^26 lognormal keys, 2^27 probes",
                 "learned_lookup_ms": c4["learned_lookup_ms"], "binary_lookup_ms": c4["binary_lookup_ms"],
                 "speedup_lookup": c4["speedup_lookup"], "speedup_join": c4["speedup_join"]}
    dist = run_nlj_dist(b)
    out["nlj_dist"] = {"workload": "distance NLJ 2^17 x 2^17 3-D points (FP64)", "ms_per_join": dist["ms_per_step"],
                       "pairs_per_s": dist["value"], "fp64_frac": dist["roofline"]["frac"]}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2",
                    choices=["c1", "c2", "c3-nlj", "c3-hidx", "c4", "c5", "nlj-dist", "stream", "copy"])
    ap.add_argument("--s-chunk", type=int, default=0, help="stream: S tuples per chunk (0: library default)")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--c5-requests", type=int, default=4096)
    ap.add_argument("--c5-lo", type=float, default=10)
    ap.add_argument("--c5-hi", type=float, default=24)
    ap.add_argument("--c5-tables", type=int, default=64)
    ap.add_argument("--c5-reuse", type=float, default=0.25)
    ap.add_argument("--c5-budget", type=float, default=200e9)
    ap.add_argument("--c5-max-batch", type=int, default=1024)
    ap.add_argument("--c5-slots", type=int, default=2)
    ap.add_argument("--c5-reps", type=int, default=3)
    ap.add_argument("--c5-qps", type=float, default=0)
    ap.add_argument("--c5-modes", default="both", choices=["both", "batched", "per_join"])
    ap.add_argument("--log2-r", type=int, default=27)
    ap.add_argument("--theta", type=float, default=0.0)
    ap.add_argument("--selectivity", type=float, default=1.0)
    ap.add_argument("--radix-bits", type=int, default=0)
    ap.add_argument("--passes", type=int, default=0)
    ap.add_argument("--lds", type=int, default=0)
    ap.add_argument("--s-tile", type=int, default=0)
    ap.add_argument("--c4-log2-r", type=int, default=26)
    ap.add_argument("--cpu-log2-r", type=int, default=27)
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--cpu-mt-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="threads of the additional N-core CPU baseline (0/1: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c5", action="store_true", help="c2: skip the batched C5 stream the default line carries")
    ap.add_argument("--no-plans", action="store_true", help="c2: skip the other partition plans (8-13 bits)")
    ap.add_argument("--no-others", action="store_true",
                    help="c2 with C5: skip the short GPU-only runs of C1, C3, C4 and the distance NLJ the default line carries")
    ap.add_argument("--host-leg", action="store_true", help="also time the host-buffer (PCIe-inclusive) path")
    ap.add_argument("--copy-all", action="store_true", help="copy: print every sweep point in the JSON line")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {a.gpus}): the library has no multi-GPU path",
              file=sys.stderr)
        return 2
    if a.config == "c2" and not a.no_c5 and a.log2_r == 27:
        # the metric's second half: the batched C5 stream (throughput, speedup, p50/p99 latency)
        res = run_c2(a)
        c5 = run_c5(a)
        res["batched"] = {"workload": c5["config"]["workload"], "tuples_per_s": c5["value"],
                          "speedup_vs_per_join": c5["speedup_batched_vs_per_join"],
                          "qps": c5["poisson"]["qps"], "p50_ms": c5["poisson"]["batched"]["p50_ms"],
                          "p99_ms": c5["poisson"]["batched"]["p99_ms"],
                          "per_join_p99_ms": c5["poisson"]["per_join"]["p99_ms"],
                          "roofline": c5["roofline"], "graph": c5["graph"]}
        if not a.no_others:
            res["other_configs"] = other_configs(a)
        print(json.dumps(res), flush=True)
        return 0
    fn = {"c1": run_c1, "c2": run_c2, "c3-nlj": run_c3_nlj, "c3-hidx": run_c3_hidx, "c4": run_c4, "c5": run_c5,
          "nlj-dist": run_nlj_dist, "stream": run_stream, "copy": run_copy}[a.config]
    print(json.dumps(fn(a)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
