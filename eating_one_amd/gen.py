"""Seeded synthetic join inputs for the five workload shapes.

The reference generates unseeded data (srand(time(NULL)),
joins/hashJoin_gpu/main_street/hj.cc:62-160); these generators are seeded
numpy PCG64 streams so that the GPU path and the CPU oracle see identical
inputs.  Tuples are int32 arrays (n, 2) = (key, payload = row index).

  c1: R = 2^16 keys, a permutation of [0, 2^16); S = 2^20 keys uniform in
      [0, 2^16)  -> every S tuple matches exactly once.
  c2: R = permutation of [0, nR); S keys: with probability `selectivity` an
      R key drawn by Zipf(theta) rank (rank r -> the key of R row r, R being
      a random permutation), otherwise uniform in the disjoint [nR, 2^31).
  c3: R, S keys uniform in [0, key_range) (duplicates on both sides).
  c4: R = nR unique lognormal(0, sigma) keys scaled into [0, 2^31); S =
      80 % keys sampled from R, 20 % uniform keys not in R.
"""
from __future__ import annotations

import numpy as np

INT31 = (1 << 31) - 1


def _tuples(keys: np.ndarray) -> np.ndarray:
    out = np.empty((len(keys), 2), dtype=np.int32)
    out[:, 0] = keys
    out[:, 1] = np.arange(len(keys), dtype=np.int32)
    return out


def zipf_ranks(rng: np.random.Generator, n_items: int, theta: float, size: int) -> np.ndarray:
    """Ranks in [0, n_items) with P(rank k) proportional to 1/(k+1)^theta.

    theta == 0 is uniform.  Otherwise rejection-inversion sampling (Hoermann &
    Derflinger, "Rejection-inversion to generate variates from monotone
    discrete distributions", ACM TOMACS 6(3), 1996), vectorised.
    """
    if size == 0:
        return np.empty(0, dtype=np.int64)
    if theta == 0.0:
        return rng.integers(0, n_items, size=size, dtype=np.int64)
    s = float(theta)
    one_m_s = 1.0 - s

    def helper1(x):  # log1p(x)/x, continuous at 0
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(np.abs(x) > 1e-8, np.log1p(x) / np.where(x == 0, 1, x), 1.0 - x * (0.5 - x / 3.0))
        return r

    def helper2(x):  # expm1(x)/x, continuous at 0
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(np.abs(x) > 1e-8, np.expm1(x) / np.where(x == 0, 1, x), 1.0 + x * 0.5 * (1.0 + x / 3.0))
        return r

    def h(x):
        return np.exp(-s * np.log(x))

    def h_int(x):
        lx = np.log(x)
        return helper2(one_m_s * lx) * lx

    def h_int_inv(x):
        t = x * one_m_s
        t = np.maximum(t, -1.0)
        return np.exp(helper1(t) * x)

    h_x1 = float(h_int(np.array(1.5))) - 1.0
    h_n = float(h_int(np.array(n_items + 0.5)))
    s_const = 2.0 - float(h_int_inv(np.array(float(h_int(np.array(2.5))) - float(h(np.array(2.0))))))

    out = np.empty(size, dtype=np.int64)
    todo = np.arange(size)
    while len(todo):
        u = h_n + rng.random(len(todo)) * (h_x1 - h_n)
        x = h_int_inv(u)
        k = np.clip(np.floor(x + 0.5), 1, n_items)
        ok = (k - x <= s_const) | (u >= h_int(k + 0.5) - h(k))
        out[todo[ok]] = k[ok].astype(np.int64) - 1
        todo = todo[~ok]
    return out


def c1(seed: int = 0xE01C1, nR: int = 1 << 16, nS: int = 1 << 20):
    rng = np.random.default_rng(seed)
    R = _tuples(rng.permutation(nR).astype(np.int32))
    S = _tuples(rng.integers(0, nR, size=nS, dtype=np.int32))
    return R, S


def c2(seed: int = 0xE01C2, nR: int = 1 << 27, nS: int = 1 << 28, theta: float = 0.0, selectivity: float = 1.0):
    rng = np.random.default_rng(seed)
    rkeys = rng.permutation(nR).astype(np.int32)
    R = _tuples(rkeys)
    hit = rng.random(nS) < selectivity if selectivity < 1.0 else np.ones(nS, dtype=bool)
    nhit = int(hit.sum())
    skeys = np.empty(nS, dtype=np.int32)
    skeys[hit] = rkeys[zipf_ranks(rng, nR, theta, nhit)]
    skeys[~hit] = rng.integers(nR, INT31, size=nS - nhit, dtype=np.int32)
    return R, _tuples(skeys)


def c3(seed: int = 0xE01C3, nR: int = 1 << 18, nS: int = 1 << 18, key_range: int = 1 << 16):
    rng = np.random.default_rng(seed)
    R = _tuples(rng.integers(0, key_range, size=nR, dtype=np.int32))
    S = _tuples(rng.integers(0, key_range, size=nS, dtype=np.int32))
    return R, S


def c4(seed: int = 0xE01C4, nR: int = 1 << 26, nS: int = 1 << 27, sigma: float = 2.0, hit_frac: float = 0.8):
    """Sorted-index workload: nR unique lognormal(0, sigma) keys scaled into [0, 2^31).

    Draws are scaled so that the 1 - 1e-6 quantile lands near 2^31, sorted, and
    made unique in one pass by k_i = cummax(d_i - i) + i: a key that collides
    moves up to the next free integer (lognormal(0, 2) packs much of its mass
    into a few hundred thousand small integers, which this fills densely).
    S: hit_frac of the probes are keys sampled from R, the rest uniform keys in
    [0, 2^31) that are not in R; shuffled."""
    rng = np.random.default_rng(seed)
    scale = INT31 / np.exp(sigma * 4.753424)
    draws = np.floor(rng.lognormal(0.0, sigma, size=nR) * scale)
    draws = np.clip(draws, 0, INT31 - nR).astype(np.int64)
    draws.sort()
    idx = np.arange(nR, dtype=np.int64)
    keys = np.maximum.accumulate(draws - idx) + idx
    rkeys = keys.astype(np.int32)
    R = _tuples(rng.permutation(rkeys))
    nhit = int(round(nS * hit_frac))
    skeys = np.empty(nS, dtype=np.int32)
    skeys[:nhit] = rkeys[rng.integers(0, nR, size=nhit)]
    miss = np.empty(0, dtype=np.int64)
    while len(miss) < nS - nhit:
        cand = rng.integers(0, INT31, size=int((nS - nhit) * 1.1) + 16, dtype=np.int64)
        pos = np.minimum(np.searchsorted(keys, cand), nR - 1)
        miss = np.concatenate([miss, cand[keys[pos] != cand]])
    skeys[nhit:] = miss[: nS - nhit].astype(np.int32)
    skeys = rng.permutation(skeys)
    return R, _tuples(skeys)


def uniform(seed: int, nR: int, nS: int, key_range: int, lo: int = 0):
    rng = np.random.default_rng(seed)
    R = _tuples(rng.integers(lo, lo + key_range, size=nR, dtype=np.int64).astype(np.int32))
    S = _tuples(rng.integers(lo, lo + key_range, size=nS, dtype=np.int64).astype(np.int32))
    return R, S


class C5Stream:
    """The C5 request stream, generated on the device (eo_gen_build /
    eo_gen_probe) and held in three device buffers: R of the requests without
    a resident table, S, and one output slot of |S| pairs per request.

    Per request i: |R|, |S| = 2^U with U ~ U[lo, hi]; with probability `reuse`
    the build side is one of `n_tables` resident tables (eo_table); R keys are
    a seeded permutation of [0, |R|); S keys hit R with probability sel ~
    U[0.1, 1] and otherwise lie in a disjoint range, so the join count of
    request i is exactly the number of its S keys below |R|."""

    def __init__(self, ctx, n_req=4096, lo=10.0, hi=24.0, n_tables=64, reuse=0.25, seed=0xE01C5):
        import eating_one_amd as eo

        rng = np.random.default_rng(seed)
        self.ctx, self.n = ctx, n_req
        t_sizes = (2 ** rng.uniform(lo, hi, size=n_tables)).astype(np.int64)
        self.tables = []  # (|R|, eo_table, the DeviceArray it was built from)
        for k, nR in enumerate(t_sizes):
            d = eo.DeviceArray(ctx, 8 * int(nR))
            ctx.gen_build(d, int(nR), seed=0x7AB1E000 + k)
            self.tables.append((int(nR), ctx.table_create(d, n=int(nR)), d))
        self.use_table = rng.random(n_req) < reuse
        self.table_id = rng.integers(0, n_tables, size=n_req)
        nR = (2 ** rng.uniform(lo, hi, size=n_req)).astype(np.int64)
        self.nR = np.where(self.use_table, t_sizes[self.table_id], nR)
        self.nS = (2 ** rng.uniform(lo, hi, size=n_req)).astype(np.int64)
        self.sel = rng.uniform(0.1, 1.0, size=n_req)
        pad = lambda n: (n + 1) // 2 * 2  # 16-byte aligned starts
        self.r_off = np.concatenate([[0], np.cumsum(np.where(self.use_table, 0, pad(self.nR)))])
        self.s_off = np.concatenate([[0], np.cumsum(pad(self.nS))])
        self.dR = eo.DeviceArray(ctx, 8 * max(int(self.r_off[-1]), 1))
        self.dS = eo.DeviceArray(ctx, 8 * max(int(self.s_off[-1]), 1))
        self.dO = eo.DeviceArray(ctx, 8 * max(int(self.s_off[-1]), 1))
        self.reqs = []
        for i in range(n_req):
            if self.use_table[i]:
                t, Rp = self.tables[int(self.table_id[i])][1], None
            else:
                t, Rp = None, self.dR.ptr + 8 * int(self.r_off[i])
                ctx.gen_build(self.dR, int(self.nR[i]), seed=0xB0000000 + i, offset=8 * int(self.r_off[i]))
            ctx.gen_probe(self.dS, int(self.nS[i]), int(self.nR[i]), float(self.sel[i]), seed=0x50000000 + i,
                          offset=8 * int(self.s_off[i]))
            self.reqs.append(eo.request(R=Rp, nR=int(self.nR[i]), S=self.dS.ptr + 8 * int(self.s_off[i]),
                                        nS=int(self.nS[i]), out=self.dO.ptr + 8 * int(self.s_off[i]),
                                        out_cap=int(self.nS[i]), table=t, tag=i))

    def _read(self, arr, off_tuples, n, cols=2):
        import ctypes

        out = np.empty((n, cols), dtype=np.int32)
        if n:
            st = self.ctx.lib.eo_copy_d2h(self.ctx.handle, out.ctypes.data, arr.ptr + 8 * int(off_tuples), out.nbytes)
            if st != 0:
                raise RuntimeError(f"eo_copy_d2h failed ({st})")
        return out

    def R(self, i):
        """Build side of request i (host copy)."""
        if self.use_table[i]:
            nR, _, d = self.tables[int(self.table_id[i])]
            return self._read(d, 0, nR)
        return self._read(self.dR, self.r_off[i], int(self.nR[i]))

    def S(self, i):
        return self._read(self.dS, self.s_off[i], int(self.nS[i]))

    def out(self, i, n):
        return self._read(self.dO, self.s_off[i], n)

    def free(self):
        for _, t, d in self.tables:
            t.free()
            d.close()
        for d in (self.dO, self.dS, self.dR):
            d.close()
        self.tables = []
