"""ctypes declarations of include/eojoin.h (libeojoin.so)."""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libeojoin.so")

P = ctypes.c_void_p
U64 = ctypes.c_uint64
U32 = ctypes.c_uint32
I32 = ctypes.c_int32
INT = ctypes.c_int
PU64 = ctypes.POINTER(ctypes.c_uint64)

EO_OK, EO_EINVAL, EO_ENOMEM, EO_ECAPACITY, EO_EHIP, EO_EBUDGET = 0, 1, 2, 3, 4, 5
EO_ENODEV, EO_EAGAIN, EO_ECANCELED, EO_ETIMEDOUT, EO_EINTERNAL = 6, 7, 8, 9, 10
EO_HOST, EO_DEVICE = 0, 1
EO_F_NO_SYNC_STATS, EO_F_DETERMINISTIC = 1, 2
EO_F_RESULT16 = 8
EO_F_LEFT_OUTER = 16
EO_F_EXACT_PARTITIONS = 32
EO_NULL_PAYLOAD = -(1 << 31)


EO_MAX_KERNEL_STATS = 12


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("ms", ctypes.c_double), ("launches", U32), ("reserved", U32)]


class JoinStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("ms_partition", ctypes.c_double),
        ("ms_count", ctypes.c_double),
        ("ms_write", ctypes.c_double),
        ("ms_build", ctypes.c_double),
        ("n_out", U64),
        ("radix_bits", U32),
        ("passes", U32),
        ("tiles", U32),
        ("lds_tuples", U32),
        ("n_kernels", U32),
        ("sampled", U32),
        ("kernels", KernelStat * EO_MAX_KERNEL_STATS),
    ]

    def kernel_table(self) -> dict:
        """{kernel family: (total ms, launches)} of the last call."""
        return {self.kernels[i].name.decode(): (self.kernels[i].ms, self.kernels[i].launches)
                for i in range(self.n_kernels)}


class JoinOpts(ctypes.Structure):
    _fields_ = [
        ("struct_size", U32),
        ("where", I32),
        ("radix_bits", I32),
        ("passes", I32),
        ("lds_tuples", I32),
        ("s_tile", I32),
        ("flags", I32),
        ("reserved", I32),
        ("stream", P),
        ("stats", ctypes.POINTER(JoinStats)),
    ]


class CtxOpts(ctypes.Structure):
    _fields_ = [
        ("struct_size", U32),
        ("flags", U32),
        ("arena_bytes", U64),
        ("hbm_budget", U64),
    ]


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("cus", I32),
        ("xcds", I32),
        ("wave_size", I32),
        ("clock_khz", I32),
        ("lds_per_wg", U64),
        ("hbm_bytes", U64),
        ("l2_bytes", U64),
    ]


class IndexInfo(ctypes.Structure):
    _fields_ = [
        ("kind", I32),
        ("learned", I32),
        ("n", U64),
        ("bytes", U64),
        ("leaves", I32),
        ("max_err", I32),
        ("mean_err", ctypes.c_double),
        ("build_ms", ctypes.c_double),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("R", P),
        ("nR", U64),
        ("table", P),
        ("S", P),
        ("nS", U64),
        ("out", P),
        ("out_cap", U64),
        ("tag", U64),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("tag", U64),
        ("n_out", U64),
        ("status", I32),
        ("reserved", U32),
    ]


class ServeOpts(ctypes.Structure):
    _fields_ = [
        ("struct_size", U32),
        ("mode", I32),
        ("max_batch", U32),
        ("slots", U32),
        ("hbm_budget", U64),
        ("deadline_ms", ctypes.POINTER(ctypes.c_double)),
        ("cancel", ctypes.POINTER(U32)),
    ]


EO_PROBE_COPY, EO_PROBE_READ, EO_PROBE_WRITE, EO_PROBE_SCATTER = 0, 1, 2, 3
EO_PROBE_NT_LOAD, EO_PROBE_NT_STORE = 1, 2
EO_VALU_COMPARE, EO_VALU_ADD = 0, 1


class ProbeOpts(ctypes.Structure):
    _fields_ = [
        ("struct_size", U32),
        ("mode", I32),
        ("bytes", U64),
        ("unroll", I32),
        ("threads", I32),
        ("blocks_per_cu", I32),
        ("flags", U32),
        ("iters", I32),
        ("run_bytes", I32),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("wall_ms", ctypes.c_double),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("tuples", U64),
        ("n_out", U64),
        ("batches", U32),
        ("largest_batch", U32),
        ("peak_admitted", U64),
        ("refused", U32),
        ("expired", U32),
        ("cancelled", U32),
        ("reserved", U32),
    ]


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("radix_bits", I32),
        ("passes", I32),
        ("lds_tuples", I32),
        ("s_tile", I32),
        ("ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
        ("candidates", U32),
        ("reserved", U32),
    ]


class StreamStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("h2d_bytes", U64),
        ("d2h_bytes", U64),
        ("n_out", U64),
        ("r_tiles", U32),
        ("s_chunks", U32),
        ("regrown", U32),
        ("pinned", U32),
    ]


class StreamOpts(ctypes.Structure):
    _fields_ = [
        ("struct_size", U32),
        ("op", I32),
        ("r_tile", U64),
        ("s_chunk", U64),
        ("out_chunk", U64),
        ("flags", I32),
        ("reserved", I32),
        ("join", ctypes.POINTER(JoinOpts)),
        ("stats", ctypes.POINTER(StreamStats)),
    ]


EO_STREAM_HASH, EO_STREAM_NLJ = 0, 1
EO_SERVE_BATCHED, EO_SERVE_PER_JOIN = 0, 1
EO_INDEX_HASH, EO_INDEX_SORTED = 0, 1
EO_LOOKUP_DEFAULT, EO_LOOKUP_BINARY, EO_LOOKUP_LEARNED = 0, 1, 2
EO_F_BINARY_SEARCH = 4
EO_REF_FILE_RAW = 1
EO_NOALLOC_SCRATCH_BYTES = 64 << 20

# name -> (restype, argtypes)
SIGNATURES = {
    "eo_abi_version": (U32, []),
    "eo_build_source_hash": (ctypes.c_char_p, []),
    "eo_status_string": (ctypes.c_char_p, [INT]),
    "eo_last_error": (ctypes.c_char_p, []),
    "eo_ctx_create": (INT, [ctypes.POINTER(P), INT, ctypes.POINTER(CtxOpts)]),
    "eo_ctx_destroy": (INT, [P]),
    "eo_ctx_sync": (INT, [P]),
    "eo_device_query": (INT, [P, ctypes.POINTER(DeviceInfo)]),
    "eo_dev_alloc": (INT, [P, U64, ctypes.POINTER(P)]),
    "eo_dev_free": (INT, [P, P]),
    "eo_host_alloc": (INT, [P, U64, ctypes.POINTER(P)]),
    "eo_host_free": (INT, [P, P]),
    "eo_copy_h2d": (INT, [P, P, P, U64]),
    "eo_copy_d2h": (INT, [P, P, P, U64]),
    "eo_copy_d2d": (INT, [P, P, P, U64]),
    "eo_bandwidth_probe": (INT, [P, U64, INT, ctypes.POINTER(ctypes.c_double)]),
    "eo_bandwidth_probe_ex": (INT, [P, ctypes.POINTER(ProbeOpts), ctypes.POINTER(ctypes.c_double)]),
    "eo_valu_probe": (INT, [P, INT, INT, INT, INT, ctypes.POINTER(ctypes.c_double)]),
    "eo_exclusive_scan_u32": (INT, [P, P, P, U64, PU64, ctypes.POINTER(JoinOpts)]),
    "eo_exclusive_scan_u64": (INT, [P, P, P, U64, PU64, ctypes.POINTER(JoinOpts)]),
    "eo_radix_partition": (INT, [P, P, U64, INT, INT, P, P, ctypes.POINTER(JoinOpts)]),
    "eo_hash_join": (INT, [P, P, U64, P, U64, P, U64, PU64, ctypes.POINTER(JoinOpts)]),
    "eo_hash_join_noalloc": (INT, [P, P, U64, P, U64, P, U64, PU64, ctypes.POINTER(JoinOpts)]),
    "eo_nlj_join": (INT, [P, P, U64, P, U64, P, U64, PU64, ctypes.POINTER(JoinOpts)]),
    "eo_nlj_dist_join": (INT, [P, P, U64, P, U64, ctypes.c_double, P, U64, PU64, ctypes.POINTER(JoinOpts)]),
    "eo_stream_join": (INT, [P, P, U64, P, U64, P, U64, PU64, ctypes.POINTER(StreamOpts)]),
    "eo_hidx_build": (INT, [P, P, U64, ctypes.POINTER(JoinOpts), ctypes.POINTER(P)]),
    "eo_sidx_build": (INT, [P, P, U64, INT, ctypes.POINTER(JoinOpts), ctypes.POINTER(P)]),
    "eo_index_join": (INT, [P, P, P, U64, P, U64, PU64, ctypes.POINTER(JoinOpts)]),
    "eo_index_lookup": (INT, [P, P, P, U64, INT, P, ctypes.POINTER(JoinOpts)]),
    "eo_index_export": (INT, [P, P, P, P]),
    "eo_index_query": (INT, [P, ctypes.POINTER(IndexInfo)]),
    "eo_index_free": (INT, [P]),
    "eo_sidx_import": (INT, [P, P, U64, INT, ctypes.POINTER(JoinOpts), ctypes.POINTER(P)]),
    "eo_ref_table_read": (INT, [ctypes.c_char_p, INT, P, U64, PU64]),
    "eo_ref_table_write": (INT, [ctypes.c_char_p, INT, P, U64]),
    "eo_ref_index_read": (INT, [ctypes.c_char_p, P, U64, PU64]),
    "eo_ref_index_write": (INT, [ctypes.c_char_p, P, U64]),
    "eo_table_create": (INT, [P, P, U64, ctypes.POINTER(JoinOpts), ctypes.POINTER(P)]),
    "eo_table_query": (INT, [P, PU64, ctypes.POINTER(I32), PU64]),
    "eo_table_free": (INT, [P]),
    "eo_batch_create": (INT, [P, ctypes.POINTER(P)]),
    "eo_batch_free": (INT, [P]),
    "eo_batch_footprint": (INT, [P, ctypes.POINTER(Request), U32, PU64]),
    "eo_batch_submit": (INT, [P, ctypes.POINTER(Request), U32, ctypes.POINTER(JoinOpts)]),
    "eo_batch_poll": (INT, [P]),
    "eo_batch_wait": (INT, [P, ctypes.POINTER(Result), U32]),
    "eo_batch_replay": (INT, [P]),
    "eo_serve_run": (INT, [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double), U32,
                           ctypes.POINTER(ServeOpts), ctypes.POINTER(Result), ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ServeStats)]),
    "eo_tune_hash_join": (INT, [P, P, U64, P, U64, P, U64, INT, ctypes.POINTER(TuneResult)]),
    "eo_tune_lookup": (INT, [P, U64, U64, ctypes.POINTER(TuneResult)]),
    "eo_tune_clear": (INT, [P]),
    "eo_tune_save": (INT, [P, ctypes.c_char_p]),
    "eo_tune_load": (INT, [P, ctypes.c_char_p]),
    "eo_gen_build": (INT, [P, P, U64, U64, ctypes.POINTER(JoinOpts)]),
    "eo_gen_probe": (INT, [P, P, U64, U64, ctypes.c_double, U64, ctypes.POINTER(JoinOpts)]),
}

_lib = None


def load():
    """Loads libeojoin.so; raises if it is missing (there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m eating_one_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
