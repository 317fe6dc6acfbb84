"""Re-encode the reference's golden Debye fixture as tests/golden/test1888.npz.

Source files (data only): test_resaxs/test1888.{pos_matrix,q_vec,ff_db,Iq_vec}.txt
in the reference tree.  Their layout is the one the reference harness parses
(test_resaxs/SAXSTest.cpp:69-151): bodies "x y z type" per line, one q per
line, a nested-bracket factor list of n_q rows x n_types values, and "q I(q)"
per line.  The npz holds the same numbers:

    pos    float32 [N, 3]
    types  int32   [N]
    q      float32 [n_q]
    table  float32 [n_q, n_types]   (row-major [q][type], as the reference)
    Iq     float64 [n_q]            golden I(q) as printed in the fixture

Run from the repository root:  python tests/golden/make_test1888.py [ref_root]
"""
import os
import re
import sys

import numpy as np


def main(ref_root="/root/reference"):
    base = os.path.join(ref_root, "test_resaxs", "test1888")
    rows = [ln.split() for ln in open(base + ".pos_matrix.txt") if ln.strip()]
    pos = np.array([[float(r[0]), float(r[1]), float(r[2])] for r in rows], dtype=np.float32)
    types = np.array([int(r[3]) for r in rows], dtype=np.int32)
    q = np.array([float(t) for t in open(base + ".q_vec.txt").read().split()], dtype=np.float32)
    text = open(base + ".ff_db.txt").read()
    vals = [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", text)]
    n_types = len(vals) // len(q)
    assert n_types * len(q) == len(vals), "factor table is not n_q x n_types"
    table = np.array(vals, dtype=np.float32).reshape(len(q), n_types)
    iq_rows = [ln.split() for ln in open(base + ".Iq_vec.txt") if ln.strip()]
    Iq = np.array([float(r[1]) for r in iq_rows], dtype=np.float64)
    assert len(Iq) == len(q)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "test1888.npz")
    np.savez_compressed(out, pos=pos, types=types, q=q, table=table, Iq=Iq)
    print(f"wrote {out}: N={len(pos)} n_q={len(q)} n_types={n_types}")


if __name__ == "__main__":
    main(*sys.argv[1:])
