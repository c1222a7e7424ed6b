"""Hand-built edge-case molecules for the parity tests (units bohr).

Each case isolates one behaviour of the reference's path:
one-centre (A = B = C), collinear atoms, ECP radial powers n = 0, 1, 2,
far-apart atoms (screened), wide exponent ranges, a minimal ECP (L = 1),
two ECP centres with different L, an atom without basis shells, and a
molecule without any ECP.
"""
from __future__ import annotations

import numpy as np

from libecp_amd.workloads import Molecule, basis_norm


def _mol(geom, basis, ecp):
    """basis: per atom list of (l, [(a, c)]); ecp: per atom None or
    list of (l, [(n, d, a)]) with the local channel the highest l."""
    shellsBS, lBS, KBS, dBS, aBS = [], [], [], [], []
    for shells in basis:
        shellsBS.append(len(shells))
        for l, prims in shells:
            lBS.append(l)
            KBS.append(len(prims))
            for a, c in prims:
                aBS.append(a)
                dBS.append(basis_norm(a, l) * c)
    shellsECP, lECP, KECP, nECP, dECP, aECP = [], [], [], [], [], []
    for e in ecp:
        if not e:
            shellsECP.append(0)
            continue
        shellsECP.append(len(e))
        for l, prims in e:
            lECP.append(l)
            KECP.append(len(prims))
            for n, d, a in prims:
                nECP.append(n)
                dECP.append(d)
                aECP.append(a)
    i32 = lambda x: np.asarray(x, dtype=np.int32)
    f64 = lambda x: np.asarray(x, dtype=np.float64)
    return Molecule(geometry=f64(geom).reshape(-1), shellsECP=i32(shellsECP), lECP=i32(lECP),
                    KECP=i32(KECP), nECP=f64(nECP), dECP=f64(dECP), aECP=f64(aECP),
                    shellsBS=i32(shellsBS), lBS=i32(lBS), KBS=i32(KBS), dBS=f64(dBS), aBS=f64(aBS))


SPD = [(0, [(12.0, 0.4), (2.1, 0.6), (0.35, 0.5)]),
       (1, [(6.0, 0.5), (1.1, 0.7)]),
       (2, [(3.0, 0.8), (0.6, 0.4)])]
ECP3 = [(0, [(2, 14.0, 6.8), (2, 3.5, 1.2)]),
        (1, [(2, -2.7, 7.9), (2, 1.1, 0.9)]),
        (2, [(2, -0.6, 6.0)]),
        (3, [(2, 0.9, 1.0), (2, -1.5, 4.0)])]


def one_centre():
    return _mol([0.0, 0.0, 0.0], [SPD], [ECP3])


def collinear():
    return _mol([0, 0, -1.7, 0, 0, 0.0, 0, 0, 2.3], [SPD, SPD, SPD], [None, ECP3, None])


def ecp_powers():
    """Radial powers n = 0, 1, 2 in every channel (local label L = 3, so the
    d shells stay inside the reference's valid domain l <= L - 1)."""
    e = [(0, [(0, 3.0, 2.5), (1, -1.2, 0.8)]),
         (1, [(1, 2.2, 1.5), (2, 0.7, 0.4)]),
         (2, [(0, -0.9, 3.0), (1, 1.4, 1.1), (2, 0.5, 0.3)]),
         (3, [(0, 0.4, 1.7), (2, -0.8, 0.9)])]
    return _mol([0.3, -0.2, 0.1, 1.4, 0.9, -1.1], [SPD[:2], SPD], [e, None])


def far_apart():
    return _mol([0, 0, 0, 25.0, 0, 0, 0, 40.0, 0], [SPD, SPD, SPD], [ECP3, None, None])


def wide_exponents():
    basis = [(0, [(250.0, 0.3), (8.0, 0.5), (0.02, 0.4)]),
             (1, [(90.0, 0.4), (0.05, 0.6)]),
             (2, [(40.0, 0.7), (0.08, 0.3)])]
    return _mol([0, 0, 0, 2.2, -1.3, 0.4, -0.8, 1.9, 2.6], [basis, basis, SPD], [ECP3, None, None])


def minimal_ecp():
    e = [(0, [(2, 5.0, 2.0)]), (1, [(2, -1.0, 1.5)])]
    return _mol([0, 0, 0, 0.0, 1.8, 0.6], [SPD[:1], SPD[:1]], [e, None])


def two_centres():
    e2 = [(0, [(2, 4.0, 3.0)]), (1, [(2, -2.0, 2.0)]), (2, [(2, 0.8, 1.0)])]
    e4 = ECP3 + [(4, [(2, 0.3, 2.0)])]
    basis_f = SPD + [(3, [(1.5, 0.9)])]
    return _mol([0, 0, 0, 2.1, 0.4, -0.7, -1.2, 1.6, 1.0], [SPD, basis_f, SPD], [e2, e4, None])


def bare_centre():
    """ECP atom without basis functions of its own."""
    return _mol([0, 0, 0, 1.3, -0.6, 0.9, -0.5, 1.7, -1.1], [[], SPD, SPD], [ECP3, None, None])


def no_ecp():
    return _mol([0, 0, 0, 1.5, 0, 0], [SPD, SPD], [None, None])


CASES = {f.__name__: f for f in (one_centre, collinear, ecp_powers, far_apart, wide_exponents,
                                 minimal_ecp, two_centres, bare_centre, no_ecp)}
