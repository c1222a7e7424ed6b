"""The CPU oracle against the reference compiled from its own sources
(oracle/_ref, present where /root/reference was available at build time):
bit-identical callback streams on seeded inputs, including derivative runs,
custom accuracy parameters and the reference's stale-buffer fallback."""
import numpy as np
import pytest

import ecp_native as E
from libecp_amd.workloads import batch_molecule, medium_cluster, plain_small
from cases import CASES

pytestmark = pytest.mark.skipif(not E.have_ref(), reason="oracle/_ref not built (reference sources absent)")


def same_stream(m, **kw):
    r1, c1 = E.ref_blocks(m, **kw)
    r2, c2 = E.oracle_blocks(m, **kw)
    assert r1 == r2
    assert c1.keys == c2.keys
    for k, a, b in zip(c1.keys, c1.blocks, c2.blocks):
        assert np.array_equal(a, b), k
    return len(c1.keys)


@pytest.mark.parametrize("seed", range(12))
def test_plain_small_seeds(seed):
    assert same_stream(plain_small(seed)) > 0


@pytest.mark.parametrize("seed", range(4))
def test_batch_recipe_molecules(seed):
    rng = np.random.default_rng(100 + seed)
    same_stream(batch_molecule(rng, int(rng.integers(3, 9))))


@pytest.mark.parametrize("n", [1, 2])
@pytest.mark.parametrize("case", ["collinear", "two_centres", "one_centre"])
def test_derivative_orders(case, n):
    same_stream(CASES[case](), n=n)


@pytest.mark.parametrize("params", [dict(tol=1e-8), dict(acc=1e-10), dict(grid=2048), dict(grid=512)])
def test_accuracy_parameters(params):
    same_stream(plain_small(3), **params)


def test_medium_cluster_full():
    assert same_stream(medium_cluster(0)) == 16002
