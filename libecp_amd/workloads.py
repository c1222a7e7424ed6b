"""Seeded synthetic ECP workloads (molecules, basis sets, ECPs).

Every generator returns a :class:`Molecule` holding the flat arrays the
reference's one-call API takes (``getIntegrals.h:7-13``): geometry is
atom-major xyz in bohr; the basis is given per atom (``shellsBS``), per shell
(``lBS``, ``KBS``) and per primitive (``dBS``, ``aBS``); the ECP likewise
(``shellsECP``; ``lECP``, ``KECP``; ``nECP``, ``dECP``, ``aECP``).  On each
ECP atom the channel with the highest l is the local one (``libecp.c:129-140``),
channels 0..L-1 are the semi-local projectors.

Recipes (all seeded, units bohr):

* ECP exponents log-uniform in [0.5, 20], coefficients N(0, sigma=5), radial
  power n = 2 (the power includes the r^2 Jacobian, ``ecp.c:54``).
* Basis exponents log-uniform (range per config), coefficients
  d = N(a, l) * c with c ~ U[0.1, 1] and
  N(a, l) = (2a/pi)^(3/4) (4a)^(l/2) / sqrt((2l-1)!!), used unnormalised by
  the integrator exactly as given.
* Every molecule respects the reference's valid input domain: the largest
  basis l is at most L-1 of every ECP and L >= 1 (``type2.c:363-364`` reads
  U_tab rows up to la+lb, which exist only for la+lb <= L-1+maxLBS).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math
import numpy as np


def _dfac(n: int) -> float:
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def basis_norm(a: float, l: int) -> float:
    return (2.0 * a / math.pi) ** 0.75 * (4.0 * a) ** (0.5 * l) / math.sqrt(_dfac(2 * l - 1))


@dataclass
class Molecule:
    geometry: np.ndarray          # float64 [3*nat]
    shellsECP: np.ndarray         # int32 [nat]
    lECP: np.ndarray              # int32 [n ecp shells]
    KECP: np.ndarray              # int32 [n ecp shells]
    nECP: np.ndarray              # float64 [n ecp primitives]
    dECP: np.ndarray
    aECP: np.ndarray
    shellsBS: np.ndarray          # int32 [nat]
    lBS: np.ndarray               # int32 [n shells]
    KBS: np.ndarray               # int32 [n shells]
    dBS: np.ndarray               # float64 [n primitives]
    aBS: np.ndarray
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nat(self) -> int:
        return int(self.shellsBS.shape[0])

    @property
    def nshells(self) -> int:
        return int(self.lBS.shape[0])

    @property
    def n_ecp_centres(self) -> int:
        return int(np.count_nonzero(self.shellsECP))

    @property
    def nao(self) -> int:
        l = self.lBS.astype(np.int64)
        return int(((l + 1) * (l + 2) // 2).sum())

    def nominal_triplets(self) -> int:
        """N_ECP * N_sh (N_sh + 1) / 2: the reference's loop count before
        screening (``libecp.c:256-312``)."""
        n = self.nshells
        return self.n_ecp_centres * n * (n + 1) // 2

    def copy(self) -> "Molecule":
        return Molecule(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                           for k, v in self.__dict__.items()})


def _positions(rng, nat, box, min_spacing):
    pts = []
    tries = 0
    while len(pts) < nat:
        p = rng.uniform(0.0, box, size=3)
        tries += 1
        if min_spacing > 0 and pts:
            d = np.sqrt(((np.asarray(pts) - p) ** 2).sum(axis=1)).min()
            if d < min_spacing:
                if tries > 200000:
                    raise RuntimeError("could not place atoms with the requested spacing")
                continue
        pts.append(p)
    return np.asarray(pts, dtype=np.float64).reshape(-1)


def build_molecule(rng, nat, box, min_spacing, ecp_atoms, L, ecp_gauss,
                   ecp_lmax_basis, other_lmax_basis, K, a_range,
                   ecp_a_range=(0.5, 20.0), ecp_sigma=5.0, ecp_n=None,
                   name=""):
    """Generic recipe: ``ecp_atoms`` is a sorted sequence of atom indices that
    carry an ECP with local channel L (projectors 0..L-1), ``ecp_gauss``
    Gaussians per channel.  ECP atoms get basis shells l = 0..ecp_lmax_basis,
    others l = 0..other_lmax_basis, each a K-primitive contraction.
    ``ecp_n``: None gives n = 2 everywhere; a sequence draws n from it."""
    geom = _positions(rng, nat, box, min_spacing)
    ecp_set = set(int(i) for i in ecp_atoms)
    shellsECP, lECP, KECP, nECP, dECP, aECP = [], [], [], [], [], []
    shellsBS, lBS, KBS, dBS, aBS = [], [], [], [], []
    lo, hi = math.log(a_range[0]), math.log(a_range[1])
    elo, ehi = math.log(ecp_a_range[0]), math.log(ecp_a_range[1])
    for i in range(nat):
        if i in ecp_set:
            shellsECP.append(L + 1)
            for l in range(L + 1):
                lECP.append(l)
                KECP.append(ecp_gauss)
                for _ in range(ecp_gauss):
                    aECP.append(math.exp(rng.uniform(elo, ehi)))
                    dECP.append(rng.normal(0.0, ecp_sigma))
                    nECP.append(2.0 if ecp_n is None else float(rng.choice(ecp_n)))
            lmax = ecp_lmax_basis
        else:
            shellsECP.append(0)
            lmax = other_lmax_basis
        shellsBS.append(lmax + 1)
        for l in range(lmax + 1):
            lBS.append(l)
            KBS.append(K)
            for _ in range(K):
                a = math.exp(rng.uniform(lo, hi))
                aBS.append(a)
                dBS.append(basis_norm(a, l) * rng.uniform(0.1, 1.0))
    i32 = lambda x: np.asarray(x, dtype=np.int32)
    f64 = lambda x: np.asarray(x, dtype=np.float64)
    return Molecule(geometry=geom, shellsECP=i32(shellsECP), lECP=i32(lECP),
                    KECP=i32(KECP), nECP=f64(nECP), dECP=f64(dECP), aECP=f64(aECP),
                    shellsBS=i32(shellsBS), lBS=i32(lBS), KBS=i32(KBS),
                    dBS=f64(dBS), aBS=f64(aBS), name=name)


def plain_small(seed=0):
    """4 atoms uniform in a 10-bohr box; atom 0 carries an ECP with local +
    s,p,d projectors (L=3), 2 Gaussians each; every atom has contracted
    Cartesian s,p,d shells of 3 primitives, exponents log-U[0.1, 50].
    40 AOs, 78 nominal shell-triplets."""
    rng = np.random.default_rng(seed)
    return build_molecule(rng, nat=4, box=10.0, min_spacing=0.0, ecp_atoms=[0],
                          L=3, ecp_gauss=2, ecp_lmax_basis=2, other_lmax_basis=2,
                          K=3, a_range=(0.1, 50.0), name=f"plain_small(seed={seed})")


def medium_cluster(seed=0, nat=50, n_ecp=10, box=30.0):
    """50-atom cluster (min spacing 2 bohr, 30-bohr box), 10 ECP centres with
    local + s..f projectors (L=4), 3 Gaussians each; ECP atoms carry s,p,d,f
    shells, the others s,p,d; 3-primitive contractions, exponents
    log-U[0.1, 50].  600 AOs, 128,800 nominal shell-triplets."""
    rng = np.random.default_rng(seed)
    ecp = np.sort(rng.choice(nat, size=n_ecp, replace=False))
    return build_molecule(rng, nat=nat, box=box, min_spacing=2.0, ecp_atoms=ecp,
                          L=4, ecp_gauss=3, ecp_lmax_basis=3, other_lmax_basis=2,
                          K=3, a_range=(0.1, 50.0), name=f"medium_cluster(seed={seed})")


def high_l(seed=0, nat=20, n_ecp=8, box=15.0):
    """20 atoms (15-bohr box, min spacing 2), 8 ECP centres with local +
    s..g projectors (L=5), 3 Gaussians each; every atom has s..g shells of 4
    primitives, exponents log-U[0.05, 100].  700 AOs, 40,400 triplets."""
    rng = np.random.default_rng(seed)
    ecp = np.sort(rng.choice(nat, size=n_ecp, replace=False))
    return build_molecule(rng, nat=nat, box=box, min_spacing=2.0, ecp_atoms=ecp,
                          L=5, ecp_gauss=3, ecp_lmax_basis=4, other_lmax_basis=4,
                          K=4, a_range=(0.05, 100.0), name=f"high_l(seed={seed})")


def batch_molecule(rng, nat):
    """One molecule of the uneven-batch / serving recipe: n_ECP =
    max(1, round(0.2 nat)) ECP atoms (L=4, 3 Gaussians per channel), basis
    l <= 3 on ECP atoms and l <= 2 elsewhere, K=3, box side
    30 (nat/50)^(1/3) with min spacing 2 bohr (relaxed for tiny boxes)."""
    n_ecp = max(1, int(round(0.2 * nat)))
    box = 30.0 * (nat / 50.0) ** (1.0 / 3.0)
    ecp = np.sort(rng.choice(nat, size=n_ecp, replace=False))
    spacing = 2.0 if box > 6.0 else 1.5
    return build_molecule(rng, nat=nat, box=box, min_spacing=spacing, ecp_atoms=ecp,
                          L=4, ecp_gauss=3, ecp_lmax_basis=3, other_lmax_basis=2,
                          K=3, a_range=(0.1, 50.0))


def uneven_batch(seed=0, n_mol=256, nat_range=(3, 40)):
    """256 independent molecules, atom counts uniform in [3, 40], 20% ECP
    atoms; basis and projectors l <= 3."""
    rng = np.random.default_rng(seed)
    mols = []
    for i in range(n_mol):
        nat = int(rng.integers(nat_range[0], nat_range[1] + 1))
        m = batch_molecule(rng, nat)
        m.name = f"uneven_batch(seed={seed})[{i}]"
        mols.append(m)
    return mols


def perturb(rng, mol, sigma=0.05):
    m = mol.copy()
    m.geometry = m.geometry + rng.normal(0.0, sigma, size=m.geometry.shape)
    return m


def serving_requests(seed=0, n=2000, nat_range=(5, 120), frac_perturbed=0.3, sigma=0.05):
    """Request stream of the serving workload: molecules of the batch recipe
    with atom counts uniform in [5, 120]; a fraction ``frac_perturbed`` of the
    requests are an earlier request's geometry displaced by N(0, sigma^2)
    bohr per coordinate.  Returns [(molecule, index of the perturbed
    original or -1)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if out and rng.random() < frac_perturbed:
            j = int(rng.integers(len(out)))
            m = perturb(rng, out[j][0], sigma)
            m.name = f"serving(seed={seed})[{i}] = perturbed [{j}]"
            out.append((m, j))
        else:
            m = batch_molecule(rng, int(rng.integers(nat_range[0], nat_range[1] + 1)))
            m.name = f"serving(seed={seed})[{i}]"
            out.append((m, -1))
    return out


CONFIGS = {
    "plain_small": plain_small,
    "medium_cluster": medium_cluster,
    "high_l": high_l,
}
