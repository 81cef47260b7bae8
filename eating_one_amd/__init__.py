"""eating_one_amd — MI355X-native inner joins (host mirror of the C ABI).

The reference drives its joins from per-program ``join()`` functions over
global tables (joins/hashJoin_gpu/main_street/hj.cc:171, joins/blockJoin_gpu/
main_street/nestLoopJoin.cc:180, joins/innerJoin_gpu/*/inner-idx-bnlj.cc:183).
Here :class:`Context` exposes the same operators over explicit arrays:
R (build side, reference "right") and S (probe side, reference "left") in,
(R payload, S payload) pairs out.  Every call goes through libeojoin.so;
there is no CPU fallback — a missing library raises ImportError.

Tuples are int32 arrays of shape (n, 2): column 0 the join key (reference
TUPLE.val), column 1 the payload / row id (reference TUPLE.key).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _abi
from ._abi import EO_DEVICE, EO_F_DETERMINISTIC, EO_HOST, JoinOpts, JoinStats

__all__ = ["Context", "DeviceArray", "Index", "EOError", "CapacityError", "load_library"]


class EOError(RuntimeError):
    def __init__(self, status: int, where: str, message: str = ""):
        self.status = status
        L = _abi.load()
        text = L.eo_status_string(status).decode()
        super().__init__(f"{where}: {text} (status {status}){': ' + message if message else ''}")


class CapacityError(EOError):
    def __init__(self, status, where, count):
        super().__init__(status, where, f"{count} matches")
        self.count = count


def load_library():
    return _abi.load()


def _check(st: int, where: str):
    if st != _abi.EO_OK:
        msg = _abi.load().eo_last_error().decode(errors="replace")
        raise EOError(st, where, msg)


def _tuples(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.int32)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError("tuples must be an int32 array of shape (n, 2)")
    return a


def _vp(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


class DeviceArray:
    """A device allocation made through eo_dev_alloc (freed on close/GC)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(ctx.lib.eo_dev_alloc(ctx.handle, self.nbytes, ctypes.byref(p)), "eo_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx: "Context", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.nbytes)
        if a.nbytes:
            _check(ctx.lib.eo_copy_h2d(ctx.handle, d.ptr, _vp(a), a.nbytes), "eo_copy_h2d")
        return d

    def to_host(self, dtype, shape, nbytes=None) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        n = out.nbytes if nbytes is None else nbytes
        if n:
            _check(self.ctx.lib.eo_copy_d2h(self.ctx.handle, _vp(out), self.ptr, n), "eo_copy_d2h")
        return out

    def close(self):
        if self.ptr is not None and self.ctx.handle:
            self.ctx.lib.eo_dev_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostArray:
    """Pinned, device-mapped host memory (eo_host_alloc): kernels can write into it directly."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(ctx.lib.eo_host_alloc(ctx.handle, self.nbytes, ctypes.byref(p)), "eo_host_alloc")
        self.ptr = p.value

    def view(self, dtype, shape) -> np.ndarray:
        """A numpy view of the memory (valid until close)."""
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        buf = (ctypes.c_char * n).from_address(self.ptr)
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    def close(self):
        if self.ptr is not None and self.ctx.handle:
            self.ctx.lib.eo_host_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One GPU context (eo_ctx): stream, scratch arena, device queries."""

    def __init__(self, device: int = 0, arena_bytes: int = 0, hbm_budget: int = 0):
        self.lib = _abi.load()
        h = ctypes.c_void_p()
        o = _abi.CtxOpts(ctypes.sizeof(_abi.CtxOpts), 0, arena_bytes, hbm_budget)
        _check(self.lib.eo_ctx_create(ctypes.byref(h), device, ctypes.byref(o)), "eo_ctx_create")
        self.handle = h.value

    def close(self):
        if self.handle:
            self.lib.eo_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(self.lib.eo_ctx_sync(self.handle), "eo_ctx_sync")

    def device_info(self) -> dict:
        info = _abi.DeviceInfo()
        _check(self.lib.eo_device_query(self.handle, ctypes.byref(info)), "eo_device_query")
        return {
            "name": info.name.decode(), "arch": info.arch.decode(), "cus": info.cus, "xcds": info.xcds,
            "wave_size": info.wave_size, "clock_khz": info.clock_khz, "lds_per_wg": info.lds_per_wg,
            "hbm_bytes": info.hbm_bytes, "l2_bytes": info.l2_bytes,
        }

    # ---- options -----------------------------------------------------------
    @staticmethod
    def opts(where=EO_HOST, radix_bits=0, passes=0, lds_tuples=0, s_tile=0, stats: JoinStats | None = None,
             deterministic: bool = False, flags: int = 0, result16: bool = False, left_outer: bool = False):
        o = JoinOpts()
        o.struct_size = ctypes.sizeof(JoinOpts)
        o.where = where
        o.radix_bits = radix_bits
        o.passes = passes
        o.lds_tuples = lds_tuples
        o.s_tile = s_tile
        o.flags = (flags | (EO_F_DETERMINISTIC if deterministic else 0) | (_abi.EO_F_RESULT16 if result16 else 0)
                   | (_abi.EO_F_LEFT_OUTER if left_outer else 0))
        o.stats = ctypes.pointer(stats) if stats is not None else None
        return o

    # ---- primitives --------------------------------------------------------
    def exclusive_scan(self, a: np.ndarray):
        """presum(): exclusive prefix sum (uint32 or uint64), returns (out, total)."""
        a = np.ascontiguousarray(a)
        out = np.empty_like(a)
        tot = ctypes.c_uint64()
        fn = {np.dtype(np.uint32): self.lib.eo_exclusive_scan_u32,
              np.dtype(np.uint64): self.lib.eo_exclusive_scan_u64}.get(a.dtype)
        if fn is None:
            raise TypeError("uint32 or uint64 array expected")
        o = self.opts()
        _check(fn(self.handle, _vp(a), _vp(out), len(a), ctypes.byref(tot), ctypes.byref(o)), "exclusive_scan")
        return out, tot.value

    def radix_partition(self, T: np.ndarray, bits: int, shift: int = 0):
        """Stable radix partition by ((uint32)key >> shift) & (2^bits-1)."""
        T = _tuples(T)
        out = np.empty_like(T)
        off = np.empty((1 << bits) + 1, dtype=np.uint64)
        o = self.opts()
        _check(self.lib.eo_radix_partition(self.handle, _vp(T), len(T), bits, shift, _vp(out), _vp(off),
                                           ctypes.byref(o)), "eo_radix_partition")
        return out, off

    # ---- joins ---------------------------------------------------------------
    def _join(self, fn, name, R, S, count_only, o):
        R, S = _tuples(R), _tuples(S)
        n = ctypes.c_uint64()
        _check(fn(self.handle, _vp(R), len(R), _vp(S), len(S), None, 0, ctypes.byref(n), ctypes.byref(o)), name)
        if count_only:
            return n.value
        cols = 4 if (o.flags & _abi.EO_F_RESULT16) else 2  # 16-byte RESULT rows or 8-byte pairs
        out = np.empty((max(n.value, 1), cols), dtype=np.int32)
        n2 = ctypes.c_uint64()
        st = fn(self.handle, _vp(R), len(R), _vp(S), len(S), _vp(out), n.value, ctypes.byref(n2), ctypes.byref(o))
        if st == _abi.EO_ECAPACITY:
            raise CapacityError(st, name, n2.value)
        _check(st, name)
        assert n2.value == n.value
        return out[: n.value]

    def hash_join(self, R, S, count_only=False, **kw):
        """Radix hash join (reference hashJoin_gpu join())."""
        return self._join(self.lib.eo_hash_join, "eo_hash_join", R, S, count_only, self.opts(**kw))

    def nlj_join(self, R, S, count_only=False, **kw):
        """Block nested-loop join (reference blockJoin_gpu join())."""
        return self._join(self.lib.eo_nlj_join, "eo_nlj_join", R, S, count_only, self.opts(**kw))

    POINT3 = np.dtype([("id", np.int32), ("pad", np.int32), ("v", np.float64, (3,))])

    def nlj_dist_join(self, R, S, distance: float = 1.0, count_only: bool = False, **kw):
        """Distance-predicate NLJ over 3-D points (reference CPU_JOIN/3d_cpu join()):
        R, S structured arrays of POINT3; returns (R id, S id) pairs."""
        R = np.ascontiguousarray(R, dtype=self.POINT3)
        S = np.ascontiguousarray(S, dtype=self.POINT3)
        o = self.opts(**kw)
        n = ctypes.c_uint64()
        fn = self.lib.eo_nlj_dist_join
        _check(fn(self.handle, _vp(R), len(R), _vp(S), len(S), float(distance), None, 0, ctypes.byref(n),
                  ctypes.byref(o)), "eo_nlj_dist_join")
        if count_only:
            return n.value
        out = np.empty((max(n.value, 1), 2), dtype=np.int32)
        n2 = ctypes.c_uint64()
        _check(fn(self.handle, _vp(R), len(R), _vp(S), len(S), float(distance), _vp(out), n.value, ctypes.byref(n2),
                  ctypes.byref(o)), "eo_nlj_dist_join")
        return out[: n.value]

    def stream_join(self, R, S, op: str = "hash", r_tile: int = 0, s_chunk: int = 0, out_chunk: int = 0,
                    result16: bool = False, left_outer: bool = False, count_only: bool = False, out=None,
                    out_cap: int | None = None, stats: "_abi.StreamStats | None" = None, join_opts=None):
        """Out-of-core streaming join over host-resident R and S (eo_stream_join): R in
        resident tiles of r_tile tuples, S streamed in chunks of s_chunk tuples.  R, S may be
        numpy arrays or HostArray views (pinned: copied by DMA directly).  `out` (optional): a
        preallocated (rows, 2|4) int32 array or HostArray view; then returns (n_out, status).
        Otherwise returns the rows (count_only: the count)."""
        R, S = _tuples(R), _tuples(S)
        o = _abi.StreamOpts()
        o.struct_size = ctypes.sizeof(_abi.StreamOpts)
        o.op = {"hash": _abi.EO_STREAM_HASH, "nlj": _abi.EO_STREAM_NLJ}[op]
        o.r_tile, o.s_chunk, o.out_chunk = r_tile, s_chunk, out_chunk
        o.flags = (_abi.EO_F_RESULT16 if result16 else 0) | (_abi.EO_F_LEFT_OUTER if left_outer else 0)
        o.join = ctypes.pointer(join_opts) if join_opts is not None else None
        o.stats = ctypes.pointer(stats) if stats is not None else None
        n = ctypes.c_uint64()
        fn = self.lib.eo_stream_join
        if out is not None:
            cap = len(out) if out_cap is None else out_cap
            st = fn(self.handle, _vp(R), len(R), _vp(S), len(S), _vp(out), cap, ctypes.byref(n), ctypes.byref(o))
            if st not in (_abi.EO_OK, _abi.EO_ECAPACITY):
                _check(st, "eo_stream_join")
            return n.value, st
        _check(fn(self.handle, _vp(R), len(R), _vp(S), len(S), None, 0, ctypes.byref(n), ctypes.byref(o)),
               "eo_stream_join")
        if count_only:
            return n.value
        rows = np.empty((max(n.value, 1), 4 if result16 else 2), dtype=np.int32)
        n2 = ctypes.c_uint64()
        _check(fn(self.handle, _vp(R), len(R), _vp(S), len(S), _vp(rows), n.value, ctypes.byref(n2),
                  ctypes.byref(o)), "eo_stream_join")
        assert n2.value == n.value
        return rows[: n.value]

    def hash_join_device(self, dR: DeviceArray, nR: int, dS: DeviceArray, nS: int, dout: DeviceArray | None,
                         out_cap: int, **kw):
        """Device-resident hash join; returns (n_out, status)."""
        n = ctypes.c_uint64()
        o = self.opts(where=EO_DEVICE, **kw)
        st = self.lib.eo_hash_join(self.handle, dR.ptr, nR, dS.ptr, nS, dout.ptr if dout else None, out_cap,
                                   ctypes.byref(n), ctypes.byref(o))
        if st not in (_abi.EO_OK, _abi.EO_ECAPACITY):
            _check(st, "eo_hash_join")
        return n.value, st

    # ---- resident indexes -------------------------------------------------------
    def hidx_build(self, R, **kw) -> "Index":
        """Hash index on R (reference innerJoin_gpu createIndex), resident on the device."""
        R = _tuples(R)
        h = ctypes.c_void_p()
        o = self.opts(**kw)
        _check(self.lib.eo_hidx_build(self.handle, _vp(R), len(R), ctypes.byref(o), ctypes.byref(h)), "eo_hidx_build")
        return Index(self, h.value)

    def sidx_import(self, sorted_pairs, learned: bool = False, **kw) -> "Index":
        """Sorted index from (key, payload) pairs already sorted by key (eo_sidx_import)."""
        P = _tuples(sorted_pairs)
        h = ctypes.c_void_p()
        o = self.opts(**kw)
        _check(self.lib.eo_sidx_import(self.handle, _vp(P), len(P), int(learned), ctypes.byref(o), ctypes.byref(h)),
               "eo_sidx_import")
        return Index(self, h.value)

    def sidx_build(self, R, learned: bool = False, **kw) -> "Index":
        """Sorted index on R (reference sort_index createIndex), optionally with the learned model."""
        R = _tuples(R)
        h = ctypes.c_void_p()
        o = self.opts(**kw)
        _check(self.lib.eo_sidx_build(self.handle, _vp(R), len(R), int(learned), ctypes.byref(o), ctypes.byref(h)),
               "eo_sidx_build")
        return Index(self, h.value)


    # ---- resident tables and batches ---------------------------------------------
    def table_create(self, R, **kw) -> "Table":
        """Partition R once on the device and keep it resident (eo_table_create).
        R is a host (n, 2) int32 array, or a DeviceArray with n= given."""
        h = ctypes.c_void_p()
        if isinstance(R, DeviceArray):
            n = kw.pop("n")
            o = self.opts(where=EO_DEVICE, **kw)
            _check(self.lib.eo_table_create(self.handle, R.ptr, n, ctypes.byref(o), ctypes.byref(h)), "eo_table_create")
        else:
            R = _tuples(R)
            o = self.opts(**kw)
            _check(self.lib.eo_table_create(self.handle, _vp(R), len(R), ctypes.byref(o), ctypes.byref(h)),
                   "eo_table_create")
        return Table(self, h.value)

    def batch(self) -> "Batch":
        return Batch(self)

    def serve(self, reqs, arrival_ms=None, mode=_abi.EO_SERVE_BATCHED, max_batch=0, slots=0, hbm_budget=0,
              deadline_ms=None, cancel=None):
        """Serve a request stream (eo_serve_run).  Returns (results, latency_ms, stats dict).

        deadline_ms: per-request admission deadlines (ms after arrival, <= 0: none);
        cancel: a uint32 numpy array of per-request cancel flags, which another
        thread may set while the loop runs."""
        n = len(reqs)
        arr = (_abi.Request * max(n, 1))(*reqs)
        res = (_abi.Result * max(n, 1))()
        lat = np.zeros(max(n, 1), dtype=np.float64)
        a = None
        if arrival_ms is not None:
            a_np = np.ascontiguousarray(arrival_ms, dtype=np.float64)
            a = a_np.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        o = _abi.ServeOpts(ctypes.sizeof(_abi.ServeOpts), mode, max_batch, slots, hbm_budget)
        if deadline_ms is not None:
            d_np = np.ascontiguousarray(deadline_ms, dtype=np.float64)
            if len(d_np) != n:
                raise ValueError("one deadline per request")
            o.deadline_ms = d_np.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if cancel is not None:
            if not (isinstance(cancel, np.ndarray) and cancel.dtype == np.uint32 and len(cancel) == n
                    and cancel.flags.c_contiguous):
                raise ValueError("cancel must be a contiguous uint32 array with one flag per request")
            o.cancel = cancel.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        st = _abi.ServeStats()
        _check(self.lib.eo_serve_run(self.handle, arr, a, n, ctypes.byref(o), res,
                                     lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(st)),
               "eo_serve_run")
        stats = {f: getattr(st, f) for f, _ in _abi.ServeStats._fields_}
        return [(r.tag, r.n_out, r.status) for r in res[:n]], lat[:n], stats

    # ---- autotuner ----------------------------------------------------------------
    def tune_hash_join(self, dR: DeviceArray, nR: int, dS: DeviceArray, nS: int, dout: DeviceArray | None = None,
                       out_cap: int = 0, reps: int = 3) -> dict:
        """Time candidate plans for this size class and store the fastest (eo_tune_hash_join)."""
        r = _abi.TuneResult()
        _check(self.lib.eo_tune_hash_join(self.handle, dR.ptr, nR, dS.ptr, nS, dout.ptr if dout else None, out_cap,
                                          reps, ctypes.byref(r)), "eo_tune_hash_join")
        return {f: getattr(r, f) for f, _ in _abi.TuneResult._fields_ if f != "reserved"}

    def tune_lookup(self, nR: int, nS: int):
        r = _abi.TuneResult()
        st = self.lib.eo_tune_lookup(self.handle, nR, nS, ctypes.byref(r))
        if st == _abi.EO_EAGAIN:
            return None
        _check(st, "eo_tune_lookup")
        return {f: getattr(r, f) for f, _ in _abi.TuneResult._fields_ if f != "reserved"}

    def tune_save(self, path: str):
        _check(self.lib.eo_tune_save(self.handle, os.fsencode(path)), "eo_tune_save")

    def tune_load(self, path: str):
        _check(self.lib.eo_tune_load(self.handle, os.fsencode(path)), "eo_tune_load")

    def tune_clear(self):
        _check(self.lib.eo_tune_clear(self.handle), "eo_tune_clear")

    def gen_build(self, dR: DeviceArray, n: int, seed: int, offset: int = 0):
        """Seeded build relation on the device: key = permutation of [0, n), payload = row."""
        o = self.opts(where=EO_DEVICE)
        _check(self.lib.eo_gen_build(self.handle, dR.ptr + offset, n, seed, ctypes.byref(o)), "eo_gen_build")

    def gen_probe(self, dS: DeviceArray, n: int, nR: int, selectivity: float, seed: int, offset: int = 0):
        """Seeded probe relation on the device: a fraction `selectivity` of keys in [0, nR)."""
        o = self.opts(where=EO_DEVICE)
        _check(self.lib.eo_gen_probe(self.handle, dS.ptr + offset, n, nR, selectivity, seed, ctypes.byref(o)),
               "eo_gen_probe")


class Table:
    """A resident, partitioned build relation (eo_table)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.handle = handle

    def info(self) -> dict:
        n = ctypes.c_uint64()
        bits = ctypes.c_int32()
        nbytes = ctypes.c_uint64()
        _check(self.ctx.lib.eo_table_query(self.handle, ctypes.byref(n), ctypes.byref(bits), ctypes.byref(nbytes)),
               "eo_table_query")
        return {"n": n.value, "radix_bits": bits.value, "bytes": nbytes.value}

    def free(self):
        # once its context is closed a table can no longer be freed through it;
        # its device memory then goes back when the process exits
        if self.handle and self.ctx.handle:
            self.ctx.lib.eo_table_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def request(R=None, nR=0, S=None, nS=0, out=None, out_cap=0, table=None, tag=0) -> "_abi.Request":
    """An eo_request from device pointers (ints), DeviceArrays or a Table."""
    def ptr(x):
        if x is None:
            return None
        return x.ptr if isinstance(x, DeviceArray) else int(x)
    return _abi.Request(ptr(R), int(nR), table.handle if table is not None else None, ptr(S), int(nS), ptr(out),
                        int(out_cap), int(tag))


class Batch:
    """A batch of joins run by one persistent probe kernel (eo_batch)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(ctx.lib.eo_batch_create(ctx.handle, ctypes.byref(h)), "eo_batch_create")
        self.handle = h.value
        self._n = 0

    def footprint(self, reqs) -> int:
        arr = (_abi.Request * max(len(reqs), 1))(*reqs)
        b = ctypes.c_uint64()
        _check(self.ctx.lib.eo_batch_footprint(self.ctx.handle, arr, len(reqs), ctypes.byref(b)), "eo_batch_footprint")
        return b.value

    def submit(self, reqs, stats: JoinStats | None = None, **kw):
        arr = (_abi.Request * max(len(reqs), 1))(*reqs)
        o = self.ctx.opts(where=EO_DEVICE, stats=stats, **kw)
        _check(self.ctx.lib.eo_batch_submit(self.handle, arr, len(reqs), ctypes.byref(o)), "eo_batch_submit")
        self._n = len(reqs)

    def replay(self):
        _check(self.ctx.lib.eo_batch_replay(self.handle), "eo_batch_replay")

    def poll(self) -> bool:
        st = self.ctx.lib.eo_batch_poll(self.handle)
        if st == _abi.EO_EAGAIN:
            return False
        _check(st, "eo_batch_poll")
        return True

    def wait(self):
        """Returns a list of (tag, n_out, status)."""
        res = (_abi.Result * max(self._n, 1))()
        _check(self.ctx.lib.eo_batch_wait(self.handle, res, self._n), "eo_batch_wait")
        return [(r.tag, r.n_out, r.status) for r in res[: self._n]]

    def run(self, reqs, **kw):
        self.submit(reqs, **kw)
        return self.wait()

    def free(self):
        if self.handle and self.ctx.handle:
            self.ctx.lib.eo_batch_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- the reference's table / index files (host only, no device needed) ------------
def read_ref_table(path: str, raw: bool = False) -> np.ndarray:
    """Reads a reference table file (int32 count + {row id, val} tuples; raw: no count)
    as (n, 2) int32 (key = val, payload = row id)."""
    L = _abi.load()
    n = ctypes.c_uint64()
    flags = _abi.EO_REF_FILE_RAW if raw else 0
    _check(L.eo_ref_table_read(os.fsencode(path), flags, None, 0, ctypes.byref(n)), "eo_ref_table_read")
    out = np.empty((n.value, 2), dtype=np.int32)
    _check(L.eo_ref_table_read(os.fsencode(path), flags, _vp(out), n.value, ctypes.byref(n)), "eo_ref_table_read")
    return out


def write_ref_table(path: str, T, raw: bool = False):
    T = _tuples(T)
    flags = _abi.EO_REF_FILE_RAW if raw else 0
    _check(_abi.load().eo_ref_table_write(os.fsencode(path), flags, _vp(T), len(T)), "eo_ref_table_write")


def read_ref_index(path: str) -> np.ndarray:
    """Reads a reference BUCKET index file as (n, 2) int32 (key = val, payload = adr), sorted by key."""
    L = _abi.load()
    n = ctypes.c_uint64()
    _check(L.eo_ref_index_read(os.fsencode(path), None, 0, ctypes.byref(n)), "eo_ref_index_read")
    out = np.empty((n.value, 2), dtype=np.int32)
    _check(L.eo_ref_index_read(os.fsencode(path), _vp(out), n.value, ctypes.byref(n)), "eo_ref_index_read")
    return out


def write_ref_index(path: str, sorted_pairs):
    P = _tuples(sorted_pairs)
    _check(_abi.load().eo_ref_index_write(os.fsencode(path), _vp(P), len(P)), "eo_ref_index_write")


class Index:
    """A resident index (eo_index) built on one Context."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.handle = handle

    def info(self) -> dict:
        i = _abi.IndexInfo()
        _check(self.ctx.lib.eo_index_query(self.handle, ctypes.byref(i)), "eo_index_query")
        return {"kind": i.kind, "learned": bool(i.learned), "n": i.n, "bytes": i.bytes, "leaves": i.leaves,
                "max_err": i.max_err, "mean_err": i.mean_err, "build_ms": i.build_ms}

    def join(self, S, count_only: bool = False, binary_search: bool = False, **kw):
        """Index join of S (reference inner-idx-bnlj.cc join())."""
        S = _tuples(S)
        flags = kw.pop("flags", 0) | (_abi.EO_F_BINARY_SEARCH if binary_search else 0)
        if kw.pop("result16", False):
            flags |= _abi.EO_F_RESULT16
        if kw.pop("left_outer", False):
            flags |= _abi.EO_F_LEFT_OUTER
        o = self.ctx.opts(flags=flags, **kw)
        L = self.ctx.lib
        n = ctypes.c_uint64()
        _check(L.eo_index_join(self.ctx.handle, self.handle, _vp(S), len(S), None, 0, ctypes.byref(n),
                               ctypes.byref(o)), "eo_index_join")
        if count_only:
            return n.value
        cols = 4 if (o.flags & _abi.EO_F_RESULT16) else 2  # 16-byte RESULT rows or 8-byte pairs
        out = np.empty((max(n.value, 1), cols), dtype=np.int32)
        n2 = ctypes.c_uint64()
        st = L.eo_index_join(self.ctx.handle, self.handle, _vp(S), len(S), _vp(out), n.value, ctypes.byref(n2),
                             ctypes.byref(o))
        _check(st, "eo_index_join")
        return out[: n.value]

    def lookup(self, S, method: int = _abi.EO_LOOKUP_DEFAULT) -> np.ndarray:
        """Per probe, the index position its search starts from (lower bound for sorted indexes)."""
        S = _tuples(S)
        pos = np.empty(len(S), dtype=np.uint64)
        o = self.ctx.opts()
        _check(self.ctx.lib.eo_index_lookup(self.ctx.handle, self.handle, _vp(S), len(S), method, _vp(pos),
                                            ctypes.byref(o)), "eo_index_lookup")
        return pos

    def export(self):
        n = self.info()["n"]
        keys = np.empty(n, dtype=np.int32)
        pays = np.empty(n, dtype=np.int32)
        _check(self.ctx.lib.eo_index_export(self.ctx.handle, self.handle, _vp(keys), _vp(pays)), "eo_index_export")
        return keys, pays

    def free(self):
        if self.handle and self.ctx.handle:
            self.ctx.lib.eo_index_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
