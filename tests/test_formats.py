"""The reference's on-disk table and index files (host-only code, no GPU).

The raw tuple layout is pinned by the reference's own joins/join_ori/R and S
files: their bytes are regenerated from glibc rand() and checked against the
committed SHA-256 digests, then read and written back through the library.
"""
import hashlib
import os

import numpy as np
import pytest

import eating_one_amd as eo
from common import join_ori_doc, libc_rand_table, ref_to_eo


@pytest.mark.parametrize("table", ["R", "S"])
def test_join_ori_files_round_trip(tmp_path, table):
    doc = join_ori_doc()["tables"][table]
    ref = libc_rand_table(doc["rows"])          # reference layout {id, val}
    raw = ref.tobytes()
    assert hashlib.sha256(raw).hexdigest() == doc["sha256"]
    path = tmp_path / table
    path.write_bytes(raw)
    got = eo.read_ref_table(str(path), raw=True)
    assert np.array_equal(got, ref_to_eo(ref))   # key = val, payload = id
    out = tmp_path / (table + ".out")
    eo.write_ref_table(str(out), got, raw=True)
    assert hashlib.sha256(out.read_bytes()).hexdigest() == doc["sha256"]


def test_table_file_with_count_header(tmp_path):
    ref = libc_rand_table(1000)
    t = ref_to_eo(ref)
    path = tmp_path / "right"
    eo.write_ref_table(str(path), t)
    b = path.read_bytes()
    assert len(b) == 4 + 8 * 1000
    assert np.frombuffer(b[:4], dtype=np.int32)[0] == 1000
    assert b[4:] == ref.tobytes()
    assert np.array_equal(eo.read_ref_table(str(path)), t)
    # empty table
    eo.write_ref_table(str(path), np.empty((0, 2), np.int32))
    assert path.read_bytes() == b"\x00\x00\x00\x00"
    assert len(eo.read_ref_table(str(path))) == 0


def test_table_file_errors(tmp_path):
    with pytest.raises(eo.EOError):
        eo.read_ref_table(str(tmp_path / "missing"))
    bad = tmp_path / "bad"
    bad.write_bytes(np.array([5], np.int32).tobytes() + b"\x00" * 16)  # says 5 rows, holds 2
    with pytest.raises(eo.EOError):
        eo.read_ref_table(str(bad))
    bad.write_bytes(b"\x00" * 12)
    with pytest.raises(eo.EOError):
        eo.read_ref_table(str(bad), raw=True)                          # not a multiple of 8


def test_index_file_round_trip(tmp_path):
    """BUCKET {val, adr} sorted by val, as the reference's createIndex writes it."""
    ref = libc_rand_table(5000)
    bucket = np.stack([ref[:, 1], np.arange(len(ref), dtype=np.int32)], 1)
    bucket = bucket[np.argsort(bucket[:, 0], kind="stable")]
    path = tmp_path / "index"
    eo.write_ref_index(str(path), bucket)
    assert path.read_bytes() == bucket.tobytes()                       # no header, fields in BUCKET order
    assert np.array_equal(eo.read_ref_index(str(path)), bucket)
