"""CPU-side checks of the C ABI: the library loads, exports every symbol
include/*.h declares, and its host-only logic behaves — no compute call."""
import ctypes
import glob
import os
import re

import pytest

from eating_one_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(eo_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("eo_hash_join", "eo_nlj_join", "eo_radix_partition", "eo_exclusive_scan_u32", "eo_ctx_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_signatures_cover_header():
    names = set(declared_symbols())
    assert set(_abi.SIGNATURES) <= names


def test_library_built_from_the_tree_sources():
    """The library embeds a hash of the sources it was built from; a stale build fails here."""
    from eating_one_amd import build as eob

    L = _abi.load()
    assert L.eo_build_source_hash().decode() == eob.source_hash(), "libeojoin.so is stale: rebuild it"


def test_build_rejects_missing_sources(monkeypatch):
    from eating_one_amd import build as eob

    monkeypatch.setattr(eob, "SOURCES", eob.SOURCES + ["no_such_file.hip"])
    with pytest.raises(RuntimeError, match="missing"):
        eob.build()


def test_status_strings_and_version():
    L = _abi.load()
    assert L.eo_abi_version() == 1
    for st in range(11):
        s = L.eo_status_string(st).decode()
        assert s and s != "unknown status"
    assert L.eo_status_string(99).decode() == "unknown status"


def test_struct_sizes_match_c_layout():
    # eo_join_opts: 8 x int32 + 2 pointers
    assert ctypes.sizeof(_abi.JoinOpts) == 8 * 4 + 2 * 8
    assert ctypes.sizeof(_abi.CtxOpts) == 24
    assert ctypes.sizeof(_abi.DeviceInfo) == 64 + 32 + 4 * 4 + 3 * 8


def test_ctx_create_without_gpu_reports_nodev():
    """On a machine without a GPU, context creation fails with EO_ENODEV (no exit, no crash)."""
    L = _abi.load()
    h = ctypes.c_void_p()
    st = L.eo_ctx_create(ctypes.byref(h), 0, None)
    if st == _abi.EO_OK:
        L.eo_ctx_destroy(h)
        pytest.skip("a GPU is visible")
    assert st == _abi.EO_ENODEV
    assert h.value is None
    assert L.eo_ctx_destroy(None) == _abi.EO_EINVAL


def test_invalid_arguments_rejected_before_device_use():
    L = _abi.load()
    n = ctypes.c_uint64()
    assert L.eo_hash_join(None, None, 0, None, 0, None, 0, ctypes.byref(n), None) == _abi.EO_EINVAL
    assert L.eo_nlj_join(None, None, 0, None, 0, None, 0, ctypes.byref(n), None) == _abi.EO_EINVAL
    assert L.eo_stream_join(None, None, 0, None, 0, None, 0, ctypes.byref(n), None) == _abi.EO_EINVAL


def test_struct_layouts_match_the_header(tmp_path):
    """Compile the header with the host C compiler and compare every struct size with ctypes."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no host C compiler")
    src = tmp_path / "sizes.c"
    src.write_text('''#include <stdio.h>
#include "eojoin.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(eo_join_opts),
         sizeof(eo_join_stats), sizeof(eo_kernel_stat), sizeof(eo_ctx_opts), sizeof(eo_device_info),
         sizeof(eo_index_info), sizeof(eo_request), sizeof(eo_result), sizeof(eo_stream_opts),
         sizeof(eo_stream_stats), sizeof(eo_serve_opts), sizeof(eo_serve_stats), sizeof(eo_probe_opts),
         sizeof(eo_tune_result), sizeof(eo_tuple));
  return 0;
}
''')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(t) for t in (_abi.JoinOpts, _abi.JoinStats, _abi.KernelStat, _abi.CtxOpts,
                                       _abi.DeviceInfo, _abi.IndexInfo, _abi.Request, _abi.Result,
                                       _abi.StreamOpts, _abi.StreamStats, _abi.ServeOpts, _abi.ServeStats,
                                       _abi.ProbeOpts, _abi.TuneResult)] + [8]
    assert got == want


def test_cli_usage_and_no_gpu_error():
    """The eo_join driver keeps the reference argv order (<op> <|S|> <|R|>)."""
    import subprocess
    exe = os.path.join(ROOT, "eating_one_amd", "eo_join")
    if not os.path.exists(exe):
        pytest.skip("eo_join not built")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "<|S|> <|R|>" in r.stderr
    r = subprocess.run([exe, "bogus", "10", "10"], capture_output=True, text=True)
    assert r.returncode == 2


def test_bench_refuses_more_than_one_gpu():
    """The bench measures one GPU: --gpus above 1 exits non-zero with a message
    before touching the device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
