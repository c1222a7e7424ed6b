"""ctypes bindings used by the tests and the benchmark.

Three libraries expose the same C interface (the reference's ``libecp.h`` /
``getIntegrals.h``):

* the product library ``libecp_amd/libecp_amd.so`` (HIP kernels behind the
  reference API) -- see :mod:`libecp_amd.native`;
* ``oracle/liboracle.so``: the CPU restatement of the reference algorithm
  (names prefixed ``orc_``), a checker only;
* ``oracle/_ref/libecp_ref.so``: the reference itself, compiled from its own
  sources by ``oracle/Makefile`` (present only where those sources were
  available at build time).
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libecp_ref.so")
REF_O0_SO = os.path.join(ROOT, "oracle", "_ref", "libecp_ref_o0.so")

CB = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                 C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_void_p)

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def _P(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def ijk_dim(l):
    return (l + 1) * (l + 2) // 2


def have_ref():
    return os.path.exists(REF_SO)


_libs = {}


def load(path):
    if path not in _libs:
        _libs[path] = C.CDLL(path)
    return _libs[path]


def _init_args(m):
    """Arguments 2..13 of libECP_init in the definition's order
    (libecp.c:54-61: lECP before KECP)."""
    return [_P(m.geometry, C.c_double), _P(m.shellsECP, C.c_int), _P(m.lECP, C.c_int),
            _P(m.KECP, C.c_int), _P(m.nECP, C.c_double), _P(m.dECP, C.c_double),
            _P(m.aECP, C.c_double), _P(m.shellsBS, C.c_int), _P(m.lBS, C.c_int),
            _P(m.KBS, C.c_int), _P(m.dBS, C.c_double), _P(m.aBS, C.c_double)]


def _gi_args(m):
    """Arguments 2..13 of getIntegrals (getIntegrals.h:7-13: KECP before lECP)."""
    return [_P(m.geometry, C.c_double), _P(m.shellsECP, C.c_int), _P(m.KECP, C.c_int),
            _P(m.lECP, C.c_int), _P(m.nECP, C.c_double), _P(m.dECP, C.c_double),
            _P(m.aECP, C.c_double), _P(m.shellsBS, C.c_int), _P(m.lBS, C.c_int),
            _P(m.KBS, C.c_int), _P(m.dBS, C.c_double), _P(m.aBS, C.c_double)]


def block_size(la, shifta, lb, shiftb, n):
    """Elements of one callback block: IJK(la+shifta) x IJK(lb+shiftb), except
    that for n = 2 the (1,0) and (0,1) shifts use the unshifted l
    (libecp.c:364-369)."""
    las, lbs = la + shifta, lb + shiftb
    if n == 2 and (shifta, shiftb) == (1, 0):
        las = la
    elif n == 2 and (shifta, shiftb) == (0, 1):
        lbs = lb
    return ijk_dim(las) * ijk_dim(lbs)


class Blocks:
    """Collects the callback stream (libecp.c:372 argument order)."""

    def __init__(self, n=0):
        self.keys = []
        self.blocks = []

        def cb(A, sa, la, shifta, B, sb, lb, shiftb, Cc, I, args):
            nn = block_size(la, shifta, lb, shiftb, n)
            self.keys.append((A, sa, la, shifta, B, sb, lb, shiftb, Cc))
            self.blocks.append(np.ctypeslib.as_array(I, shape=(nn,)).copy())
        self.cb = CB(cb)


def ref_getIntegrals(m, grid=1024, tol=1e-12, acc=1e-14, lib=None):
    lib = lib or load(REF_SO)
    nao = m.nao
    I = np.zeros((nao, nao))
    rc = lib.getIntegrals(C.c_int(m.nat), *_gi_args(m), C.c_int(grid), C.c_double(tol),
                          C.c_double(acc), C.c_int(nao), _P(I, C.c_double))
    return rc, I


def ref_blocks(m, n=0, grid=1024, tol=1e-12, acc=1e-14, lib=None):
    lib = lib or load(REF_SO)
    lib.libECP_init.restype = C.c_void_p
    h = lib.libECP_init(C.c_int(m.nat), *_init_args(m), C.c_int(n), C.c_int(-1), None,
                        C.c_int(grid), C.c_double(tol), C.c_double(acc))
    if not h:
        return None, None
    col = Blocks(n)
    rc = lib.calculateECPIntegrals(C.c_void_p(h), col.cb, None)
    lib.libECP_free(C.c_void_p(h))
    return rc, col


def oracle_getIntegrals(m, grid=1024, tol=1e-12, acc=1e-14, options=0):
    lib = load(ORACLE_SO)
    nao = m.nao
    I = np.zeros((nao, nao))
    rc = lib.orc_getIntegrals(C.c_int(m.nat), *_gi_args(m), C.c_int(grid), C.c_double(tol),
                              C.c_double(acc), C.c_int(nao), _P(I, C.c_double), C.c_int(options))
    return rc, I


class OracleStats(C.Structure):
    _fields_ = [("ps93_calls", C.c_long * 2), ("ps93_fail", C.c_long * 2),
                ("psm92_calls", C.c_long * 2), ("psm92_fail", C.c_long * 2),
                ("left_breaks", C.c_long * 2), ("blocks", C.c_long * 2),
                ("triplets_evaluated", C.c_long), ("fb_tabulations", C.c_long * 2),
                ("fb_points", C.c_long * 2), ("ps93_level", (C.c_long * 16) * 2),
                ("psm92_level", (C.c_long * 11) * 2)]


def oracle_blocks(m, n=0, grid=1024, tol=1e-12, acc=1e-14, options=0, stats=False):
    lib = load(ORACLE_SO)
    lib.orc_init.restype = C.c_void_p
    h = lib.orc_init(C.c_int(m.nat), *_init_args(m), C.c_int(n), C.c_int(-1), None,
                     C.c_int(grid), C.c_double(tol), C.c_double(acc), C.c_int(options))
    if not h:
        return (None, None, None) if stats else (None, None)
    col = Blocks(n)
    rc = lib.orc_calculate(C.c_void_p(h), col.cb, None)
    st = OracleStats()
    lib.orc_get_stats(C.c_void_p(h), C.byref(st))
    lib.orc_free(C.c_void_p(h))
    return (rc, col, st) if stats else (rc, col)
