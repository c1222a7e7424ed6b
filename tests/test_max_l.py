"""The largest angular momenta this build evaluates: basis l plus derivative
order up to 6, so type-1 classes up to lambda = la + lb = 12 (the kernels'
largest template) and Bessel orders up to L - 1 + 6 = 12, with L = 7
(semi-local channels l = 0..6, the local channel l = 7), inside the
reference's valid domain (basis l + n <= L - 1).  The reference sizes its
tables from the inputs (libecp.c:129-149, 176-185) and has no fixed cap.

Four molecules: i shells (l = 6) at n = 0 (classes up to lambda = 12), h
shells (l = 5) at n = 1 (one shell shifted to l = 6: lambda up to 11), and
under L = 6 h shells at n = 0 (lambda up to 10) and g shells at n = 1
(lambda up to 9).  On the CPU the oracle is bit-identical to the reference
on all four (where oracle/_ref is built); on the device every block is
within 1e-10 absolute + 1e-8 relative of the oracle, widened by the
reference's own rounding floor at these l (2 |ref - ref_fma|, ref_fma the
oracle compiled with FMA contraction).
"""
import ctypes as C
import os

import numpy as np
import pytest

import cases
import ecp_native as E
from libecp_amd import native

GEOM = [0.0, 0.0, 0.0, 1.6, 0.4, -0.7, -0.9, 1.3, 0.6]
ECP6 = [(0, [(2, 9.0, 3.1), (2, 1.6, 0.8)]), (1, [(2, -3.0, 2.7)]), (2, [(2, 1.2, 1.9)]),
        (3, [(2, -0.8, 1.4)]), (4, [(2, 0.5, 1.1)]), (5, [(2, -0.3, 0.9)]), (6, [(2, 0.2, 1.3)])]

ECP7 = ECP6 + [(7, [(2, 0.15, 1.2)])]

def _shell(l, *prims):
    return (l, list(prims))


def h_shells():
    basis = [[_shell(0, (4.0, 0.5), (0.6, 0.6)), _shell(2, (1.8, 0.7)), _shell(5, (1.2, 0.6), (0.5, 0.5))],
             [_shell(1, (2.2, 0.6)), _shell(3, (1.1, 0.8)), _shell(5, (0.9, 0.7))],
             [_shell(0, (1.5, 0.9)), _shell(4, (0.8, 0.6))]]
    return cases._mol(GEOM, basis, [ECP6, None, ECP6])


def g_shells():
    basis = [[_shell(0, (4.0, 0.5), (0.6, 0.6)), _shell(4, (1.3, 0.6), (0.5, 0.5))],
             [_shell(2, (1.6, 0.7)), _shell(4, (0.9, 0.7))],
             [_shell(1, (1.5, 0.9)), _shell(3, (0.8, 0.6))]]
    return cases._mol(GEOM, basis, [ECP6, None, ECP6])


def i_shells():
    basis = [[_shell(0, (4.0, 0.5), (0.6, 0.6)), _shell(3, (1.8, 0.7)), _shell(6, (1.1, 0.6), (0.5, 0.5))],
             [_shell(1, (2.2, 0.6)), _shell(4, (1.0, 0.8)), _shell(6, (0.8, 0.7))],
             [_shell(2, (1.5, 0.9)), _shell(5, (0.7, 0.6))]]
    return cases._mol(GEOM, basis, [ECP7, None, ECP7])


def h_shells_l7():
    basis = [[_shell(0, (4.0, 0.5), (0.6, 0.6)), _shell(5, (1.2, 0.6), (0.5, 0.5))],
             [_shell(2, (1.6, 0.7)), _shell(5, (0.9, 0.7))],
             [_shell(1, (1.5, 0.9)), _shell(3, (0.8, 0.6))]]
    return cases._mol(GEOM, basis, [ECP7, None, ECP7])


MOLS = {"h_shells_n0": (h_shells, 0, 10), "g_shells_n1": (g_shells, 1, 9),
        "i_shells_n0": (i_shells, 0, 12), "h_shells_n1": (h_shells_l7, 1, 11)}


@pytest.mark.skipif(not E.have_ref(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("name", sorted(MOLS))
def test_oracle_equals_reference_at_max_l(name):
    make, n, lam = MOLS[name]
    m = make()
    r1, c1 = E.ref_blocks(m, n=n)
    r2, c2 = E.oracle_blocks(m, n=n)
    assert r1 == r2 == 0
    assert c1.keys == c2.keys and len(c1.keys) > 0
    assert max(k[2] + k[3] + k[6] + k[7] for k in c1.keys) == lam
    for k, a, b in zip(c1.keys, c1.blocks, c2.blocks):
        assert np.array_equal(a, b), k


def _fma_blocks(m, n):
    lib = C.CDLL(os.path.join(E.ROOT, "oracle", "liboracle_fma.so"))
    lib.orc_init.restype = C.c_void_p
    h = lib.orc_init(C.c_int(m.nat), *E._init_args(m), C.c_int(n), C.c_int(-1), None, C.c_int(1024),
                     C.c_double(1e-12), C.c_double(1e-14), C.c_int(1))
    col = E.Blocks(n)
    rc = lib.orc_calculate(C.c_void_p(h), col.cb, None)
    lib.orc_free(C.c_void_p(h))
    return rc, col


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MOLS))
def test_device_at_max_l(name):
    make, n, lam = MOLS[name]
    m = make()
    h = native.Handle(m, n=n)
    try:
        rc, keys, blocks = h.blocks()
    finally:
        h.close()
    rc2, col = E.oracle_blocks(m, n=n, options=1)
    rc3, fma = _fma_blocks(m, n)
    assert rc == rc2 == rc3 == 0
    assert keys == col.keys
    assert max(k[2] + k[3] + k[6] + k[7] for k in keys) == lam
    for k, a, b, f in zip(keys, blocks, col.blocks, fma.blocks):
        a, b, f = np.asarray(a), np.asarray(b), np.asarray(f)
        assert np.all(np.abs(a - b) <= 1e-10 + 1e-8 * np.abs(b) + 2 * np.abs(b - f)), \
            (k, float(np.abs(a - b).max()))
