"""Density contraction of a callback stream, in numpy: the checker for the
device consumer (libecp_contract / libecp_batch_contract).

The callback stream (libecp.c:372) delivers, per (C, shell a on A, shell
b >= a on B), a type-1 and a type-2 block.  With a symmetric density P:

* n = 0: E = sum_{mu nu} P_{mu nu} V_{mu nu}.  A block of two different
  shells stands for itself and its transpose (weight 2), a block of one
  shell with itself for itself (weight 1).
* n = 1: the blocks come with shifted angular momenta; the derivative of a
  primitive is d/dA_d g(a) = 2 zeta g(a + 1_d) - a_d g(a - 1_d)
  (libecp.c:242-245), the zeta factor already in the (+1) block.  The ECP
  centre's derivative follows from translational invariance,
  d/dC = -(d/dA + d/dB): each block's contribution goes to the atom of its
  shifted shell with + and to the centre with -.

These formulas are pinned independently by finite differences of E in the
tests (the gradient of n = 1 against E(R + h) - E(R - h) of n = 0).
"""
from __future__ import annotations

import numpy as np


def ijk(l):
    return (l + 1) * (l + 2) // 2


def components(l):
    """libint Cartesian order of a shell (the default shellOrdering):
    x-power descending, then y-power descending."""
    out = []
    for i in range(l + 1):
        for j in range(i + 1):
            out.append((l - i, i - j, j))
    return out


def ao_offsets(m):
    off, o, s = {}, 0, 0
    for A in range(m.nat):
        for k in range(int(m.shellsBS[A])):
            off[(A, k)] = o
            o += ijk(int(m.lBS[s]))
            s += 1
    return off


def contract_energy(m, keys, blocks, P):
    off = ao_offsets(m)
    E = 0.0
    for key, blk in zip(keys, blocks):
        A, sa, la, sha, B, sb, lb, shb, Cc = key
        a0, b0 = off[(A, sa)], off[(B, sb)]
        V = np.asarray(blk).reshape(ijk(la), ijk(lb))
        w = 1.0 if (A, sa) == (B, sb) else 2.0
        E += w * float(np.sum(P[a0:a0 + ijk(la), b0:b0 + ijk(lb)] * V))
    return E


def contract_gradient(m, keys, blocks, P):
    off = ao_offsets(m)
    g = np.zeros((m.nat, 3))
    for key, blk in zip(keys, blocks):
        A, sa, la, sha, B, sb, lb, shb, Cc = key
        a0, b0 = off[(A, sa)], off[(B, sb)]
        w = 1.0 if (A, sa) == (B, sb) else 2.0
        V = np.asarray(blk).reshape(ijk(la + sha), ijk(lb + shb))
        Pab = P[a0:a0 + ijk(la), b0:b0 + ijk(lb)]
        bra = sha != 0
        l, sh = (la, sha) if bra else (lb, shb)
        comps = components(l)
        shifted = {c: k for k, c in enumerate(components(l + sh))}
        s = np.zeros(3)
        for i, c in enumerate(comps):
            for d in range(3):
                x = list(c)
                if sh > 0:
                    coef = 2.0
                    x[d] += 1
                else:
                    if x[d] == 0:
                        continue
                    coef = -float(x[d])
                    x[d] -= 1
                r = shifted[tuple(x)]
                if bra:
                    s[d] += coef * float(np.dot(Pab[i, :], V[r, :]))
                else:
                    s[d] += coef * float(np.dot(Pab[:, i], V[:, r]))
        X = A if bra else B
        g[X] += w * s
        g[Cc] -= w * s
    return g


def random_density(nao, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(nao, max(1, nao // 2)))
    return X @ X.T / nao
