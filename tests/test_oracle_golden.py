"""The CPU oracle against the golden fixtures written by the reference itself
(tests/golden/make_golden.py): the callback stream must be identical, bit for
bit, including the derivative orders n = 1, 2 and the getIntegrals matrix."""
import numpy as np
import pytest

import ecp_native as E
from conftest import GOLDEN, golden_ids, load_golden


@pytest.mark.parametrize("path", GOLDEN, ids=golden_ids())
def test_oracle_matches_reference_fixture(path):
    g = load_golden(path)
    rc, col = E.oracle_blocks(g["mol"], n=g["n"])
    assert rc == g["rc"]
    assert col.keys == g["keys"]
    for k, a, b in zip(col.keys, col.blocks, g["blocks"]):
        # NaN where the reference reads past its factorial table (L = 1 with
        # s-only basis: tmDim = 0, transformations.h:13 reads fac[3])
        assert np.array_equal(a, b, equal_nan=True), k
    if "matrix_upper" in g:
        rc2, I = E.oracle_getIntegrals(g["mol"])
        assert rc2 == 0
        assert np.array_equal(I[np.triu_indices(g["mol"].nao)], g["matrix_upper"], equal_nan=True)


def test_fixtures_present():
    assert len(GOLDEN) >= 10
