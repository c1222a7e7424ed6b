"""Host logic of the product library, on the CPU: exported C ABI, input
validation / error codes, and the screening that decides which triplets the
device evaluates (checked against the oracle's callback stream)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import ecp_native as E
from cases import CASES
from libecp_amd import native
from libecp_amd.workloads import medium_cluster, plain_small

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof") or name.isupper():
                continue
            prefix = text[:m.start()].rstrip()
            if prefix.endswith("(*") or prefix.endswith("typedef"):
                continue
            names.add(name)
    return sorted(n for n in names if not n.startswith("ECP") and n != "void")


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    names = declared_functions()
    assert {"libECP_init", "calculateECPIntegrals", "libECP_free", "getIntegrals"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n


def handle(m, **kw):
    lib = native.lib()
    h = lib.libECP_init(C.c_int(m.nat), *E._init_args(m), C.c_int(kw.get("n", 0)), C.c_int(kw.get("lmax", -1)),
                        kw.get("ordering"), C.c_int(kw.get("grid", 1024)), C.c_double(kw.get("tol", 1e-12)),
                        C.c_double(kw.get("acc", 1e-14)))
    return h


def test_init_rejects_derivative_order_above_two():
    assert not handle(plain_small(0), n=3)
    assert "derivative" in native.last_error()
    assert not handle(plain_small(0), n=-1)


def test_derivative_order_limited_by_basis_l():
    # basis up to d (l = 2): n = 2 reaches l = 4 and is accepted; an i shell
    # (l = 6, the largest basis l) leaves no room for a shift
    m = plain_small(0)
    keys, _ = screened(m, n=2)
    assert len(keys) > 0
    m.lBS[-1] = 6
    lib = native.lib()
    h = handle(m, n=1)
    assert h
    assert lib.libecp_screened_triplets(C.c_void_p(h), None, 0) == -1
    assert "derivative" in native.last_error()
    lib.libECP_free(C.c_void_p(h))


def test_init_rejects_local_only_ecp():
    m = plain_small(0)
    m.shellsECP[0] = 1          # only the first ECP shell: l = 0 becomes the local channel
    assert not handle(m)
    assert native.last_error()


@pytest.mark.parametrize("grid", [100, 255])
def test_init_rejects_small_large_grid(grid):
    assert not handle(plain_small(0), grid=grid)


def test_init_reports_bessel_tabulation_failure():
    """accuracy < 0: the series never meets its stopping test (bessel.c:51)."""
    assert not handle(plain_small(0), acc=-1.0)
    assert "Bessel" in native.last_error()


def test_init_rejects_short_custom_ordering():
    ordering = np.zeros(3 * 200, dtype=np.int32)
    assert not handle(plain_small(0), lmax=2, ordering=ordering.ctypes.data_as(C.POINTER(C.c_int)))


def screened(m, n=0):
    lib = native.lib()
    h = handle(m, n=n)
    assert h
    n = lib.libecp_screened_triplets(C.c_void_p(h), None, 0)
    keys = np.zeros(5 * max(n, 1), dtype=np.int32)
    lib.libecp_screened_triplets(C.c_void_p(h), keys.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(n))
    lib.libecp_nominal_triplets.restype = C.c_longlong
    nominal = lib.libecp_nominal_triplets(C.c_void_p(h))
    lib.libECP_free(C.c_void_p(h))
    return [tuple(int(v) for v in k) for k in keys[:5 * n].reshape(-1, 5)], nominal


@pytest.mark.parametrize("name", sorted(CASES))
def test_screening_matches_oracle_edge_cases(name):
    m = CASES[name]()
    keys, _ = screened(m)
    rc, col = E.oracle_blocks(m, options=1)
    expect = [(k[8], k[0], k[1], k[4], k[5]) for k in col.keys[::2]]
    assert keys == expect


@pytest.mark.parametrize("seed", range(6))
def test_screening_matches_oracle_seeds(seed):
    m = plain_small(seed)
    keys, nominal = screened(m)
    rc, col = E.oracle_blocks(m, options=1)
    assert keys == [(k[8], k[0], k[1], k[4], k[5]) for k in col.keys[::2]]
    assert nominal == m.nominal_triplets() == 78


@pytest.mark.parametrize("n", [1, 2])
@pytest.mark.parametrize("seed", range(3))
def test_derivative_screening_matches_oracle(n, seed):
    # one record per (triplet, shift); A = B = C triplets and the shifts the
    # reference skips (libecp.c:304-330) are absent
    m = plain_small(seed)
    keys, _ = screened(m, n=n)
    rc, col = E.oracle_blocks(m, n=n, options=1)
    assert keys == [(k[8], k[0], k[1], k[4], k[5]) for k in col.keys[::2]]


@pytest.mark.parametrize("name", ["collinear", "one_centre", "two_centres"])
def test_derivative_screening_matches_oracle_edge_cases(name):
    m = CASES[name]()
    for n in (1, 2):
        keys, _ = screened(m, n=n)
        rc, col = E.oracle_blocks(m, n=n, options=1)
        assert keys == [(k[8], k[0], k[1], k[4], k[5]) for k in col.keys[::2]]


def test_screening_matches_oracle_medium_cluster():
    m = medium_cluster(0)
    keys, nominal = screened(m)
    rc, col = E.oracle_blocks(m, options=1)
    assert len(keys) == 8001
    assert keys == [(k[8], k[0], k[1], k[4], k[5]) for k in col.keys[::2]]
    assert nominal == 128800


def gpu_present():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(gpu_present(), reason="checks the behaviour without a GPU")
def test_calculation_fails_loudly_without_device():
    m = plain_small(0)
    h = native.Handle(m)
    rc, keys, blocks = h.blocks()
    assert rc == -1 and keys == []
    assert "device" in native.last_error().lower() or "hip" in native.last_error().lower()
    rc, _ = native.get_integrals(m)
    assert rc == 1
