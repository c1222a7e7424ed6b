import glob
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

from libecp_amd.workloads import Molecule  # noqa: E402

FIELDS = ("geometry", "shellsECP", "lECP", "KECP", "nECP", "dECP", "aECP",
          "shellsBS", "lBS", "KBS", "dBS", "aBS")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden(path):
    z = np.load(path)  # allow_pickle stays False
    m = Molecule(**{f: z[f] for f in FIELDS}, name=os.path.basename(path))
    sizes = z["sizes"]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    blocks = [z["values"][offs[i]:offs[i + 1]] for i in range(len(sizes))]
    keys = [tuple(int(v) for v in k) for k in z["keys"]]
    out = dict(mol=m, rc=int(z["rc"]), n=int(z["n"]), keys=keys, blocks=blocks)
    if "matrix_upper" in z.files:
        out["matrix_upper"] = z["matrix_upper"]
    return out


GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))


def golden_ids():
    return [os.path.basename(p)[4:-4] for p in GOLDEN]


def within(gpu, ref, atol=1e-10, rtol=1e-8, envelope=None):
    """|gpu - ref| <= atol + rtol |ref| (+ envelope); returns (ok, worst excess)."""
    bound = atol + rtol * np.abs(ref)
    if envelope is not None:
        bound = bound + envelope
    d = np.abs(gpu - ref)
    return bool((d <= bound).all()), float((d - bound).max()) if d.size else 0.0
