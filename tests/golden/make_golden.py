"""Writes the golden fixtures of tests/golden/ by running the reference
library compiled from its own sources (oracle/_ref/libecp_ref.so, built by
oracle/Makefile from /root/reference/src).  Run in the build container:

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture (ref_<case>.npz) holds the molecule's input arrays and the
reference's outputs: the return code, the callback stream (keys of 9 ints
and the concatenated blocks, libecp.c:372 order) for derivative order n,
and, for n = 0, the getIntegrals matrix's upper triangle.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import ecp_native as E  # noqa: E402
from cases import CASES  # noqa: E402
from libecp_amd.workloads import plain_small  # noqa: E402

FIELDS = ("geometry", "shellsECP", "lECP", "KECP", "nECP", "dECP", "aECP",
          "shellsBS", "lBS", "KBS", "dBS", "aBS")


def dump(name, m, n=0, matrix=True):
    rc, col = E.ref_blocks(m, n=n)
    keys = np.asarray(col.keys, dtype=np.int32).reshape(-1, 9)
    sizes = np.asarray([b.size for b in col.blocks], dtype=np.int64)
    vals = np.concatenate(col.blocks) if col.blocks else np.zeros(0)
    out = {f: getattr(m, f) for f in FIELDS}
    out.update(rc=np.int32(rc), n=np.int32(n), keys=keys, sizes=sizes, values=vals)
    if matrix and n == 0:
        grc, I = E.ref_getIntegrals(m)
        iu = np.triu_indices(m.nao)
        out.update(matrix_rc=np.int32(grc), matrix_upper=I[iu])
    path = os.path.join(HERE, f"ref_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{path}: rc={rc} blocks={len(col.blocks)}")


def main():
    if not E.have_ref():
        sys.exit("oracle/_ref/libecp_ref.so is missing: run `make -C oracle ref` first")
    for s in (0, 1, 2, 6):
        dump(f"plain_small_{s}", plain_small(s))
    for name, f in CASES.items():
        dump(name, f())
    dump("collinear_n1", CASES["collinear"](), n=1, matrix=False)
    dump("ecp_powers_n2", CASES["ecp_powers"](), n=2, matrix=False)


if __name__ == "__main__":
    main()
