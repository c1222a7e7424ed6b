"""The CPU oracle over many molecules at once (test infrastructure).

Each molecule's getIntegrals matrix from ``oracle/liboracle.so`` (the CPU
restatement of the reference; options=1: zeros past a large-grid screening
break, the device's semantics), computed in a pool of freshly spawned
processes.  Spawned, not forked: the test process has usually initialised
the GPU runtime, which a forked child must not inherit.  The workers load
only the oracle (ctypes) and numpy.
"""
from __future__ import annotations

import multiprocessing as mp
import os


def _one(m):
    import ecp_native as E
    rc, I = E.oracle_getIntegrals(m, options=1)
    return rc, I


def _init(paths):
    import sys
    for p in paths:
        if p not in sys.path:
            sys.path.insert(0, p)


def oracle_matrices(mols, workers=None):
    """[(rc, I)] for each molecule, in order."""
    here = os.path.dirname(os.path.abspath(__file__))
    paths = [os.path.dirname(here), here]
    if workers is None:
        n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        workers = max(1, min(16, n, len(mols)))
    if workers == 1:
        _init(paths)
        return [_one(m) for m in mols]
    # largest first, so that the longest oracle runs start at once
    order = sorted(range(len(mols)), key=lambda i: -mols[i].nominal_triplets())
    with mp.get_context("spawn").Pool(workers, initializer=_init, initargs=(paths,)) as pool:
        res = pool.map(_one, [mols[i] for i in order], chunksize=1)
    out = [None] * len(mols)
    for i, r in zip(order, res):
        out[i] = r
    return out
