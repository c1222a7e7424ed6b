"""The reference's own rounding floor on the medium cluster, and the device
inside the strict bar.

The oracle is bit-identical to the reference.  Compiled with FMA
contraction (oracle/liboracle_fma.so: the same source, only the rounding
of a*b+c changes) it moves the medium cluster's getIntegrals matrix by up
to 5.2e-10 absolute (10 elements above 1e-10): ulp-level changes move the
stopping level of an adaptive quadrature, and the element moves by that
level's error.  The device library is compiled without contraction
(-ffp-contract=off), so its integrands and level sums round as the
reference's do; what remains is its own exp/pow/trigonometric rounding.
Bar for the device: every element within 1e-10 absolute of the reference
and no further from it than the FMA build.
"""
import os

import numpy as np
import pytest

import ecp_native as E
from libecp_amd import native
from libecp_amd.workloads import medium_cluster


def _matrices():
    m = medium_cluster(0)
    rc, ref = E.oracle_getIntegrals(m, options=1)
    so, E.ORACLE_SO = E.ORACLE_SO, os.path.join(E.ROOT, "oracle", "liboracle_fma.so")
    try:
        rc2, fma = E.oracle_getIntegrals(m, options=1)
    finally:
        E.ORACLE_SO = so
    assert rc == rc2 == 0
    return m, ref, fma


def test_fma_build_of_the_reference_algorithm_moves_the_matrix():
    _, ref, fma = _matrices()
    d = np.abs(fma - ref)
    assert 1e-10 < d.max() < 1e-8
    assert np.all(d <= 1e-10 + 1e-8 * np.abs(ref))


@pytest.mark.gpu
def test_device_within_1e10_of_the_reference():
    m, ref, fma = _matrices()
    rc, I = native.get_integrals(m)
    assert rc == 0
    d = np.abs(I - ref)
    assert d.max() <= np.abs(fma - ref).max(), (d.max(), np.abs(fma - ref).max())
    assert d.max() <= 1e-10, (float(d.max()), int((d > 1e-10).sum()))
