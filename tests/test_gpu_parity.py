"""Parity of the HIP library (through its C ABI) with the reference.

Bar: |gpu - ref| <= 1e-10 + 1e-8 |ref| element-wise on every callback block
and on the getIntegrals matrix.  References: the golden fixtures written by
the reference itself, and the CPU oracle (bit-identical to the reference,
with the large-grid fallback's stale-buffer reuse replaced by zeros -- the
device semantics -- which moves the reference's results by <= 2.2e-12 on
these inputs).  For the high-l configuration a few elements are decided by
last-bit rounding inside the reference's adaptive quadrature; there the bar
adds 2 |ref - ref_fma|, where ref_fma is the same oracle compiled with FMA
contraction (the reference's own rounding sensitivity)."""
import ctypes as C
import math

import numpy as np
import pytest

import ecp_native as E
from cases import CASES
from conftest import GOLDEN, golden_ids, load_golden, within
from libecp_amd import native
from libecp_amd.workloads import high_l, medium_cluster, plain_small

pytestmark = pytest.mark.gpu

ATOL, RTOL = 1e-10, 1e-8


def gpu_blocks(m, **kw):
    h = native.Handle(m, **kw)
    try:
        return h.blocks()
    finally:
        h.close()


def check_stream(keys, blocks, ref_keys, ref_blocks, envelope=None):
    assert keys == ref_keys
    worst = -1.0
    for i, (k, a, b) in enumerate(zip(keys, blocks, ref_blocks)):
        env = None if envelope is None else envelope[i]
        ok, excess = within(a, b, ATOL, RTOL, env)
        worst = max(worst, excess)
        assert ok, (k, "type", 1 + i % 2, excess)
    return worst


@pytest.mark.parametrize("path", [p for p in GOLDEN if "_n1" not in p and "_n2" not in p],
                         ids=[i for i in golden_ids() if not i.endswith(("_n1", "_n2"))])
def test_against_reference_fixtures(path):
    g = load_golden(path)
    m = g["mol"]
    rc, keys, blocks = gpu_blocks(m)
    assert rc == g["rc"]
    assert keys == g["keys"]
    finite_ref = [np.isfinite(b).all() for b in g["blocks"]]
    for i, (k, a, b) in enumerate(zip(keys, blocks, g["blocks"])):
        assert np.isfinite(a).all(), k
        if finite_ref[i]:
            ok, excess = within(a, b, ATOL, RTOL)
            assert ok, (k, 1 + i % 2, excess)
    rc, I = native.get_integrals(m)
    assert rc == 0
    up = I[np.triu_indices(m.nao)]
    fin = np.isfinite(g["matrix_upper"])
    assert within(up[fin], g["matrix_upper"][fin], ATOL, RTOL)[0]
    assert np.isfinite(up).all()


@pytest.mark.parametrize("path", [p for p in GOLDEN if "_n1" in p or "_n2" in p],
                         ids=[i for i in golden_ids() if i.endswith(("_n1", "_n2"))])
def test_derivative_fixtures(path):
    """Derivative integrals (n = 1, 2): every shifted block of the reference's
    callback stream."""
    g = load_golden(path)
    rc, keys, blocks = gpu_blocks(g["mol"], n=g["n"])
    assert rc == g["rc"]
    assert keys == g["keys"]
    for i, (k, a, b) in enumerate(zip(keys, blocks, g["blocks"])):
        if np.isfinite(b).all():
            ok, excess = within(a, b, ATOL, RTOL)
            assert ok, (k, 1 + i % 2, excess)


@pytest.mark.parametrize("n", [1, 2])
@pytest.mark.parametrize("seed", [0, 3])
def test_derivatives_against_oracle(n, seed):
    m = plain_small(seed)
    rc, keys, blocks = gpu_blocks(m, n=n)
    rc2, col = E.oracle_blocks(m, n=n, options=1)
    assert rc == rc2 == 0
    check_stream(keys, blocks, col.keys, col.blocks)


def test_derivatives_medium_cluster_against_oracle():
    m = medium_cluster(0, nat=16, n_ecp=4, box=14.0)
    rc, keys, blocks = gpu_blocks(m, n=1)
    rc2, col = E.oracle_blocks(m, n=1, options=1)
    assert rc == rc2 == 0
    check_stream(keys, blocks, col.keys, col.blocks)


def test_minimal_ecp_type1_exact():
    """L = 1 with an s basis: the reference reads past its factorial table
    (NaN); the device result is checked against the exact three-centre
    integral of the n = 2 local Gaussian instead."""
    from test_oracle_math import overlap3_1d, shells_of
    m = CASES["minimal_ecp"]()
    rc, keys, blocks = gpu_blocks(m)
    G = m.geometry.reshape(-1, 3)
    sh = shells_of(m)
    local = [(2, -1.0, 1.5)]
    for i, (k, b) in enumerate(zip(keys, blocks)):
        if i % 2:
            continue
        A, sa, B, sb = k[0], k[1], k[4], k[5]
        _, _, aa, da = sh[sum(m.shellsBS[:A]) + sa]
        _, _, ab, db = sh[sum(m.shellsBS[:B]) + sb]
        v = sum(c1 * c2 * d * overlap3_1d(0, 0, G[A][0], G[B][0], G[0][0], x1, x2, g)
                * overlap3_1d(0, 0, G[A][1], G[B][1], G[0][1], x1, x2, g)
                * overlap3_1d(0, 0, G[A][2], G[B][2], G[0][2], x1, x2, g)
                for x1, c1 in zip(aa, da) for x2, c2 in zip(ab, db) for _, d, g in local)
        assert abs(b[0] - v) <= 1e-9 * max(1.0, abs(v)), (k, b[0], v)


@pytest.mark.parametrize("seed", range(10))
def test_plain_small_against_oracle(seed):
    m = plain_small(seed)
    rc, keys, blocks = gpu_blocks(m)
    rc2, col = E.oracle_blocks(m, options=1)
    assert rc == rc2 == 0
    check_stream(keys, blocks, col.keys, col.blocks)


def test_medium_cluster_against_oracle():
    m = medium_cluster(0)
    rc, keys, blocks = gpu_blocks(m)
    rc2, col = E.oracle_blocks(m, options=1)
    assert rc == rc2 == 0
    assert len(keys) == 16002
    check_stream(keys, blocks, col.keys, col.blocks)


def _oracle_fma():
    import os
    path = os.path.join(E.ROOT, "oracle", "liboracle_fma.so")
    lib = C.CDLL(path)
    lib.orc_init.restype = C.c_void_p
    return lib


def test_high_l_against_oracle_with_rounding_envelope():
    """g shells, exponents down to 0.05: some type-1 elements are lost to
    cancellation in the reference's C-centred expansion and binomial shift
    (values off the exact integral by up to 1e2 relative); they move with the
    last bit of any operation.  Bar per block: the usual tolerance widened by
    2 |ref - ref_fma|; type-1 blocks outside it must be no further from the
    exact three-centre integral (n = 2 local Gaussians) than the reference."""
    from test_oracle_math import cart_exps, overlap3_1d, shells_of
    m = high_l(0)
    rc, keys, blocks = gpu_blocks(m)
    rc2, col = E.oracle_blocks(m, options=1)
    lib = _oracle_fma()
    h = lib.orc_init(C.c_int(m.nat), *E._init_args(m), C.c_int(0), C.c_int(-1), None, C.c_int(1024),
                     C.c_double(1e-12), C.c_double(1e-14), C.c_int(1))
    fma = E.Blocks(0)
    rc3 = lib.orc_calculate(C.c_void_p(h), fma.cb, None)
    lib.orc_free(C.c_void_p(h))
    assert rc == rc2 == rc3 == 0
    assert keys == col.keys
    G = m.geometry.reshape(-1, 3)
    sh = shells_of(m)
    local = {}
    k = p = 0
    for at, ns in enumerate(m.shellsECP):
        rows = []
        for _ in range(ns):
            for _ in range(int(m.KECP[k])):
                rows.append((int(m.lECP[k]), m.nECP[p], m.dECP[p], m.aECP[p]))
                p += 1
            k += 1
        if rows:
            L = max(r[0] for r in rows)
            local[at] = [r for r in rows if r[0] == L]
    outside = 0
    for i, (key, a, b, f) in enumerate(zip(keys, blocks, col.blocks, fma.blocks)):
        ok, _ = within(a, b, ATOL, RTOL, 2 * np.abs(b - f))
        if ok:
            continue
        assert i % 2 == 0, (key, "type-2 block outside the rounding envelope")
        outside += 1
        A, sa, B, sb, Cat = key[0], key[1], key[4], key[5], key[8]
        _, la, aa, da = sh[sum(m.shellsBS[:A]) + sa]
        _, lb, ab, db = sh[sum(m.shellsBS[:B]) + sb]
        ex = np.array([sum(c1 * c2 * d * math.prod(overlap3_1d(ea[x], eb[x], G[A][x], G[B][x], G[Cat][x], x1, x2, g)
                                                    for x in range(3))
                           for x1, c1 in zip(aa, da) for x2, c2 in zip(ab, db) for _, n, d, g in local[Cat])
                       for ea in cart_exps(la) for eb in cart_exps(lb)])
        assert np.all(np.abs(a - ex) <= 2 * np.abs(b - ex) + ATOL + RTOL * np.abs(ex)), key
    # the device rounds as the reference does (no contraction); one type-1
    # block of two g shells on one atom (A = B = 19, C = 5) stays outside
    # the envelope: its values, ~1e-5 in a block of 1e-4, come from
    # cancellation in the C-centred expansion, and it is checked against
    # the exact integral above
    assert outside <= 1


def test_failure_code_and_delivered_prefix():
    """tolerance = 0 fails every adaptive integral that is not exactly zero:
    the first failing block stops the stream with the documented code."""
    m = plain_small(1)
    h = native.Handle(m, tol=0.0)
    rc, keys, blocks = h.blocks()
    h.close()
    rc2, col = E.oracle_blocks(m, tol=0.0, options=1)
    assert rc == rc2 and rc in (1, 2)
    assert keys == col.keys
    check_stream(keys, blocks, col.keys, col.blocks)


def test_custom_shell_ordering():
    """A non-libint Cartesian ordering (reversed inside every shell) through
    libECP_init's lmax / shellOrdering arguments."""
    m = plain_small(2)
    lmax = 8
    order = []
    for l in range(lmax + 1):
        comps = [(l - i, i - j, j) for i in range(l + 1) for j in range(i + 1)]
        for c in reversed(comps):
            order.extend(c)
    ordering = np.asarray(order, dtype=np.int32)
    rc, keys, blocks = gpu_blocks(m, lmax=lmax, ordering=ordering)
    lib = E.load(E.ORACLE_SO)
    h = lib.orc_init(C.c_int(m.nat), *E._init_args(m), C.c_int(0), C.c_int(lmax),
                     ordering.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(1024), C.c_double(1e-12),
                     C.c_double(1e-14), C.c_int(1))
    col = E.Blocks(0)
    rc2 = lib.orc_calculate(C.c_void_p(h), col.cb, None)
    lib.orc_free(C.c_void_p(h))
    assert rc == rc2 == 0
    check_stream(keys, blocks, col.keys, col.blocks)


def test_repeat_runs_are_bitwise_identical():
    m = medium_cluster(1)
    h = native.Handle(m)
    r1 = h.blocks()
    r2 = h.blocks()
    h.close()
    assert r1[0] == r2[0] == 0 and r1[1] == r2[1]
    assert all(np.array_equal(a, b) for a, b in zip(r1[2], r2[2]))


def test_getintegrals_accumulates_upper_triangle():
    m = plain_small(4)
    base = np.arange(m.nao * m.nao, dtype=np.float64).reshape(m.nao, m.nao) * 1e-3
    rc, I = native.get_integrals(m, I=base.copy())
    rc0, I0 = native.get_integrals(m)
    assert rc == rc0 == 0
    # blocks are added one by one into the caller's values: equal up to rounding
    assert np.allclose(I, base + I0, rtol=1e-14, atol=1e-14)
    assert np.all(np.tril(I0, -1) == 0)
