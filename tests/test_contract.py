"""Density contraction on the device (libecp_contract, libecp_batch_contract).

The checker is tests/contract_ref.py applied to callback blocks.  Its two
formulas are pinned on the CPU, independently of the device:
* the energy equals sum(P * V) over the full symmetric matrix that
  getIntegrals accumulates (upper triangle mirrored);
* the gradient of n = 1 blocks equals central finite differences of the
  n = 0 energy (to 1e-7 relative; the difference converges as h^2) for
  molecules the reference resolves accurately: moderate exponents, shifted
  shells l + 1 <= L - 1.
The device results must then agree with the checker on the device's own
blocks to rounding (1e-12 relative) and on the oracle's blocks within the
block tolerance, the batch must equal the handle bit for bit, and the
device-pointer path must equal the host path.
"""
import ctypes

import numpy as np
import pytest

import cases
import ecp_native as E
from contract_ref import contract_energy, contract_gradient, random_density
from libecp_amd import native
from libecp_amd.workloads import batch_molecule, plain_small


SOFT = [(0, [(1.2, 0.5), (0.4, 0.5)]), (1, [(0.9, 0.7), (0.3, 0.4)]), (2, [(0.8, 0.6)])]
ECP_SOFT = [(0, [(2, 4.0, 1.8), (2, 1.5, 0.7)]), (1, [(2, -2.0, 1.5), (2, 0.8, 0.6)]), (2, [(2, -0.6, 1.2)]),
            (3, [(2, 0.9, 0.8)]), (4, [(2, 0.4, 1.1)])]
ECP3_SOFT = ECP_SOFT[:3] + [(3, [(2, 0.4, 1.1)])]
GEOM = [0, 0, 0, 1.4, 0.3, -0.5, -0.6, 1.1, 0.8]


def soft_spd():
    """d shells (shifted to f) with L = 4, two ECP centres with basis."""
    return cases._mol(GEOM, [SOFT, SOFT, SOFT[:2]], [ECP_SOFT, None, ECP_SOFT])


def soft_sp():
    return cases._mol(GEOM, [SOFT[:2], SOFT[:2], SOFT[:1]], [ECP3_SOFT, ECP3_SOFT, None])


def soft_bare_centre():
    return cases._mol([0, 0, 0, 1.2, -0.6, 0.9, -0.5, 1.3, -0.8], [[], SOFT, SOFT[:2]], [ECP_SOFT, None, None])


# Molecules whose energy the reference resolves accurately and smoothly.
# With tight primitives far from the centre its type-2 values carry
# quadrature errors at the per-cent level (a diagonal s block 2.6 bohr from
# the centre: 2 % against an independent quadrature), which no derivative
# formula reconciles with a finite difference.
SMOOTH = {"soft_spd": soft_spd, "soft_sp": soft_sp, "soft_bare_centre": soft_bare_centre}


def _oracle_energy(m, P):
    rc, col = E.oracle_blocks(m, n=0, options=1)
    assert rc == 0
    return contract_energy(m, col.keys, col.blocks, P)


@pytest.mark.parametrize("name", ["plain_small", "collinear", "ecp_powers", "two_centres"])
def test_energy_formula_equals_matrix_trace(name):
    m = plain_small(0) if name == "plain_small" else cases.CASES[name]()
    P = random_density(m.nao, 5)
    rc, I = E.oracle_getIntegrals(m, options=1)
    V = I + I.T - np.diag(np.diag(I))
    e = _oracle_energy(m, P)
    assert abs(e - float(np.sum(P * V))) <= 1e-12 * max(1.0, abs(e))


@pytest.mark.parametrize("name", sorted(SMOOTH))
def test_gradient_formula_matches_finite_differences(name):
    m = SMOOTH[name]()
    P = random_density(m.nao, 7)
    rc, col = E.oracle_blocks(m, n=1, options=1)
    assert rc == 0
    g = contract_gradient(m, col.keys, col.blocks, P)
    h = 1e-4
    fd = np.zeros_like(g)
    for X in range(m.nat):
        for d in range(3):
            mp, mm = m.copy(), m.copy()
            mp.geometry[3 * X + d] += h
            mm.geometry[3 * X + d] -= h
            fd[X, d] = (_oracle_energy(mp, P) - _oracle_energy(mm, P)) / (2 * h)
    assert np.abs(g).max() > 1e-2
    assert np.abs(g - fd).max() <= 1e-7 * np.abs(g).max(), np.abs(g - fd).max()


def test_contract_rejects_second_derivatives():
    m = soft_sp()
    h = native.Handle(m, n=2)
    try:
        with pytest.raises(RuntimeError, match="orders 0 and 1"):
            h.contract(np.eye(m.nao))
    finally:
        h.close()


# ---------------------------------------------------------------- device

GPU_CASES = ["plain_small", "collinear", "ecp_powers", "two_centres", "bare_centre", "wide_exponents",
             "soft_spd", "soft_bare_centre"]


def _case(name):
    if name == "plain_small":
        return plain_small(0)
    if name in SMOOTH:
        return SMOOTH[name]()
    return cases.CASES[name]()


def _scale(blocks, P):
    """max |P| sum |V| over the blocks: the size of the contraction's terms
    (up to the derivative coefficients, at most 2 l + 2)."""
    return float(np.abs(P).max()) * sum(float(np.abs(b).sum()) for b in blocks)


@pytest.mark.gpu
@pytest.mark.parametrize("name", GPU_CASES)
def test_device_energy(name):
    m = _case(name)
    P = random_density(m.nao, 11)
    h = native.Handle(m, n=0)
    try:
        rc, e = h.contract(P)
        rcb, keys, blocks = h.blocks()
    finally:
        h.close()
    assert rc == rcb == 0
    s = _scale(blocks, P)
    assert abs(e - contract_energy(m, keys, blocks, P)) <= 1e-12 * s
    rco, col = E.oracle_blocks(m, n=0, options=1)
    assert rco == 0
    assert abs(e - contract_energy(m, col.keys, col.blocks, P)) <= 1e-10 * s + 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", GPU_CASES)
def test_device_gradient(name):
    m = _case(name)
    P = random_density(m.nao, 13)
    h = native.Handle(m, n=1)
    try:
        rc, g = h.contract(P)
        rcb, keys, blocks = h.blocks()
    finally:
        h.close()
    assert rc == rcb == 0
    assert g.shape == (m.nat, 3)
    s = _scale(blocks, P)
    assert np.abs(g - contract_gradient(m, keys, blocks, P)).max() <= 1e-12 * s
    rco, col = E.oracle_blocks(m, n=1, options=1)
    assert rco == 0
    assert np.abs(g - contract_gradient(m, col.keys, col.blocks, P)).max() <= 1e-10 * s + 1e-10


@pytest.mark.gpu
def test_device_gradient_matches_finite_differences():
    m = soft_spd()
    P = random_density(m.nao, 17)
    h = native.Handle(m, n=1)
    try:
        g = h.contract(P)[1]
    finally:
        h.close()
    hh = 1e-4
    fd = np.zeros_like(g)
    for X in range(m.nat):
        for d in range(3):
            es = []
            for sgn in (1, -1):
                mx = m.copy()
                mx.geometry[3 * X + d] += sgn * hh
                hx = native.Handle(mx, n=0)
                try:
                    es.append(hx.contract(P)[1])
                finally:
                    hx.close()
            fd[X, d] = (es[0] - es[1]) / (2 * hh)
    assert np.abs(g - fd).max() <= 1e-7 * np.abs(g).max()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1])
def test_batch_equals_handles_and_device_pointers(n):
    # device buffers from the HIP runtime the library itself runs on
    hip = ctypes.CDLL("libamdhip64.so.7")
    rng = np.random.default_rng(21)
    mols = [batch_molecule(rng, int(k)) for k in (3, 7, 12, 5)] + [soft_spd()]
    Ps = [random_density(m.nao, 30 + i) for i, m in enumerate(mols)]
    b = native.Batch(mols, n=n)
    Pd, Od = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        b.execute(1)
        codes, outs, ms = b.contract(Ps)
        nP, nout = b.contract_sizes()
        Ph = np.concatenate([p.ravel() for p in Ps])
        Oh = np.zeros(nout)
        assert hip.hipMalloc(ctypes.byref(Pd), ctypes.c_size_t(8 * nP)) == 0
        assert hip.hipMalloc(ctypes.byref(Od), ctypes.c_size_t(8 * nout)) == 0
        assert hip.hipMemcpy(Pd, ctypes.c_void_p(Ph.ctypes.data), ctypes.c_size_t(8 * nP), 1) == 0
        codes2, ms2 = b.contract_device(Pd.value, Od.value)
        assert hip.hipMemcpy(ctypes.c_void_p(Oh.ctypes.data), Od, ctypes.c_size_t(8 * nout), 2) == 0
        outs2 = b._split(Oh)
    finally:
        b.close()
        for p in (Pd, Od):
            if p.value:
                hip.hipFree(p)
    assert list(codes) == [0] * len(mols) and list(codes2) == [0] * len(mols)
    assert ms > 0.0
    for i, m in enumerate(mols):
        h = native.Handle(m, n=n)
        try:
            rc, v = h.contract(Ps[i])
        finally:
            h.close()
        assert rc == 0
        assert np.array_equal(np.asarray(v), np.asarray(outs[i])), (i, n)
        assert np.array_equal(np.asarray(outs2[i]), np.asarray(outs[i])), (i, n)


@pytest.mark.gpu
def test_failure_code_follows_the_callback_order():
    m = plain_small(1)
    P = random_density(m.nao, 3)
    h = native.Handle(m, n=0, tol=0.0)
    try:
        rc, _ = h.contract(P)
        rcb, _, _ = h.blocks()
    finally:
        h.close()
    assert rc == rcb and rc in (1, 2)


@pytest.mark.gpu
def test_batch_matrix_refused_for_derivative_batch():
    b = native.Batch([soft_sp()], n=1)
    try:
        b.execute(1)
        b.fetch()
        rc, _ = b.matrix(0)
        assert rc == -1 and "derivative" in native.last_error()
    finally:
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1])
def test_contraction_under_custom_ordering(n):
    """A custom Cartesian ordering (reversed inside every shell, through
    libECP_init's lmax / shellOrdering) with the density permuted the same
    way gives the same energy and gradient: the contraction reads the
    components (and their shifted partners) through the handle's ordering."""
    m = plain_small(2)
    lmax = 10
    order = []
    for l in range(lmax + 1):
        comps = [(l - i, i - j, j) for i in range(l + 1) for j in range(i + 1)]
        for c in reversed(comps):
            order.extend(c)
    ordering = np.asarray(order, dtype=np.int32)
    perm, o = [], 0
    for l in m.lBS:
        nl = (int(l) + 1) * (int(l) + 2) // 2
        perm.extend(o + nl - 1 - c for c in range(nl))   # new position -> libint position
        o += nl
    perm = np.asarray(perm)
    P = random_density(m.nao, 23)
    Pr = P[np.ix_(perm, perm)]
    h0 = native.Handle(m, n=n)
    h1 = native.Handle(m, n=n, lmax=lmax, ordering=ordering)
    try:
        rc0, v0 = h0.contract(P)
        rc1, v1 = h1.contract(Pr)
    finally:
        h0.close()
        h1.close()
    assert rc0 == rc1 == 0
    v0, v1 = np.asarray(v0), np.asarray(v1)
    assert np.allclose(v0, v1, rtol=1e-12, atol=1e-12 * np.abs(v0).max()), (v0, v1)


def test_contract_rejects_short_rows():
    m = soft_sp()
    h = native.Handle(m, n=0)
    try:
        with pytest.raises(RuntimeError, match="rowdim"):
            h.contract(np.zeros((m.nao, m.nao - 1)))
    finally:
        h.close()
