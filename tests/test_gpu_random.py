"""Seeded random molecules against the oracle, through the C ABI.

Each molecule draws its own shape: 2-7 atoms, one to three ECP centres with
local labels L = 2..5 and 1-3 Gaussians per channel (radial powers 0, 1, 2),
basis shells up to l = min(L-1, 3) with contraction lengths 1-4 mixed within
the molecule, and a derivative order n = 0, 1, 2.  Every callback block must
agree with the oracle (a bit-exact restatement of the reference) to
1e-10 absolute + 1e-8 relative, and the failure code and delivered prefix
must match.
"""
import math

import numpy as np
import pytest

import ecp_native as E
from conftest import within
from libecp_amd import native
from libecp_amd.workloads import Molecule, basis_norm

pytestmark = pytest.mark.gpu

ATOL, RTOL = 1e-10, 1e-8


def random_molecule(seed):
    rng = np.random.default_rng(1000 + seed)
    nat = int(rng.integers(2, 8))
    box = 3.0 + 1.5 * nat
    geom = rng.uniform(0.0, box, size=3 * nat)
    n_ecp = int(rng.integers(1, min(3, nat) + 1))
    ecp_atoms = set(int(a) for a in rng.choice(nat, size=n_ecp, replace=False))
    L = {a: int(rng.integers(2, 6)) for a in ecp_atoms}
    lmax_basis = min(min(L.values()) - 1, 3)
    shellsECP, lECP, KECP, nECP, dECP, aECP = [], [], [], [], [], []
    shellsBS, lBS, KBS, dBS, aBS = [], [], [], [], []
    for i in range(nat):
        if i in ecp_atoms:
            shellsECP.append(L[i] + 1)
            for l in range(L[i] + 1):
                k = int(rng.integers(1, 4))
                lECP.append(l)
                KECP.append(k)
                for _ in range(k):
                    nECP.append(float(rng.choice([0.0, 1.0, 2.0])))
                    dECP.append(float(rng.normal(0.0, 5.0)))
                    aECP.append(math.exp(rng.uniform(math.log(0.5), math.log(20.0))))
        else:
            shellsECP.append(0)
        ls = sorted(int(v) for v in rng.integers(0, lmax_basis + 1, size=int(rng.integers(1, 4))))
        shellsBS.append(len(ls))
        for l in ls:
            k = int(rng.integers(1, 5))
            lBS.append(l)
            KBS.append(k)
            for _ in range(k):
                a = math.exp(rng.uniform(math.log(0.1), math.log(50.0)))
                aBS.append(a)
                dBS.append(basis_norm(a, l) * rng.uniform(0.1, 1.0))
    i32 = lambda x: np.asarray(x, dtype=np.int32)
    f64 = lambda x: np.asarray(x, dtype=np.float64)
    m = Molecule(geometry=f64(geom), shellsECP=i32(shellsECP), lECP=i32(lECP), KECP=i32(KECP),
                 nECP=f64(nECP), dECP=f64(dECP), aECP=f64(aECP), shellsBS=i32(shellsBS), lBS=i32(lBS),
                 KBS=i32(KBS), dBS=f64(dBS), aBS=f64(aBS), name=f"random({seed})")
    n = int(rng.integers(0, 3))
    if lmax_basis + n > 5:
        n = 5 - lmax_basis
    return m, n


@pytest.mark.parametrize("seed", range(16))
def test_random_molecule_against_oracle(seed):
    m, n = random_molecule(seed)
    h = native.Handle(m, n=n)
    try:
        rc, keys, blocks = h.blocks()
    finally:
        h.close()
    rc2, col = E.oracle_blocks(m, n=n, options=1)
    assert rc == rc2
    assert keys == col.keys
    for i, (k, a, b) in enumerate(zip(keys, blocks, col.blocks)):
        if not np.isfinite(b).all():
            continue
        ok, excess = within(a, b, ATOL, RTOL)
        assert ok, (m.name, n, k, 1 + i % 2, excess)


@pytest.mark.parametrize("seed", range(4))
def test_general_radial_powers_against_oracle(seed):
    """ECP radial powers outside 0, 1, 2 (0.5, 1.5, 3): the device takes
    them through its general pow, the oracle through C's pow."""
    m, n = random_molecule(seed)
    rng = np.random.default_rng(7000 + seed)
    m.nECP = rng.choice([0.5, 1.5, 2.0, 3.0], size=m.nECP.shape).astype(np.float64)
    assert np.any(m.nECP != 2.0)
    h = native.Handle(m, n=n)
    try:
        rc, keys, blocks = h.blocks()
    finally:
        h.close()
    rc2, col = E.oracle_blocks(m, n=n, options=1)
    assert rc == rc2
    assert keys == col.keys
    for i, (k, a, b) in enumerate(zip(keys, blocks, col.blocks)):
        if not np.isfinite(b).all():
            continue
        ok, excess = within(a, b, ATOL, RTOL)
        assert ok, (m.name, n, k, 1 + i % 2, excess)


def test_random_batch_equals_molecules_alone():
    mols = [random_molecule(s)[0] for s in range(12)]
    b = native.Batch(mols)
    try:
        b.execute(1)
        b.fetch()
        got = [b.matrix(i) for i in range(len(mols))]
    finally:
        b.close()
    for m, (rc, I) in zip(mols, got):
        rc1, I1 = native.get_integrals(m)
        assert rc == rc1 == 0
        assert np.array_equal(I, I1), m.name
