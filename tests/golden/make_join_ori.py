"""Writes tests/golden/join_ori.json from the reference's join_ori data files.

Run in the build container only (it reads /root/reference).  The fixtures are
the raw 8-byte {id, val} tables R (1000 rows) and S (100000 rows) of
joins/join_ori/ and their stable `val % 4` partitions hash-part-{R,S}-{0..3}
written by joins/join_ori/hj.cc:60-87.  The JSON keeps their SHA-256 digests
and sizes; the tests regenerate the tables from glibc rand() and compare.
"""
import hashlib
import json
import os
import sys

import numpy as np

REF = "/root/reference/joins/join_ori"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "join_ori.json")


def main():
    doc = {"source": "joins/join_ori (reference data files)", "tables": {}, "partitions": {}}
    for t in ("R", "S"):
        raw = open(os.path.join(REF, t), "rb").read()
        doc["tables"][t] = {"rows": len(raw) // 8, "sha256": hashlib.sha256(raw).hexdigest()}
        for p in range(4):
            praw = open(os.path.join(REF, f"hash-part-{t}-{p}"), "rb").read()
            doc["partitions"][f"{t}-{p}"] = {"rows": len(praw) // 8, "sha256": hashlib.sha256(praw).hexdigest()}
    R = np.frombuffer(open(os.path.join(REF, "R"), "rb").read(), dtype=np.int32).reshape(-1, 2)
    S = np.frombuffer(open(os.path.join(REF, "S"), "rb").read(), dtype=np.int32).reshape(-1, 2)
    # brute-force equi-join on val (column 1): count and sum of R.val over matches
    cr = np.bincount(R[:, 1], minlength=100).astype(np.int64)
    cs = np.bincount(S[:, 1], minlength=100).astype(np.int64)
    doc["join"] = {"pairs": int((cr * cs).sum()), "sum_r_val": int((cr * cs * np.arange(100)).sum())}
    json.dump(doc, open(OUT, "w"), indent=1, sort_keys=True)
    print(json.dumps(doc["join"]))


if __name__ == "__main__":
    sys.exit(main())
