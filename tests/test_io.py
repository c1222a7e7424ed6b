"""File loaders (include/libecp_io.h) and the ecp_integrals program.

The fixtures tests/golden/example_test_c.{bs,ecp} are the reference's own
example inputs (example/test_c.*), copied as data.  Their ECP records have
three columns, "coefficient exponent power", which the reference's example
program misreads with its four-column read (ex1.c:112); the loader reads
each record by its column count, so the same data written in ex1.c's four
columns must load to identical arrays.
"""
import os
import subprocess

import numpy as np
import pytest

from libecp_amd import native
from libecp_amd.build import CLI

HERE = os.path.dirname(os.path.abspath(__file__))
ECP = os.path.join(HERE, "golden", "example_test_c.ecp")
BS = os.path.join(HERE, "golden", "example_test_c.bs")


def write_xyz(path, geom, comment="two carbons"):
    g = np.asarray(geom).reshape(-1, 3)
    with open(path, "w") as f:
        f.write(f"{len(g)}\n{comment}\n")
        for x, y, z in g:
            f.write(f"C {x:.17g} {y:.17g} {z:.17g}\n")


@pytest.fixture
def xyz(tmp_path):
    p = str(tmp_path / "c2.xyz")
    write_xyz(p, [0.0, 0.0, 0.0, 0.4, -0.3, 2.6])
    return p


def test_example_files_load(xyz):
    m = native.load_molecule(xyz, ECP, BS)
    assert m.nat == 2
    assert list(m.geometry) == [0.0, 0.0, 0.0, 0.4, -0.3, 2.6]
    # test_c.ecp: per atom "6 2 4", channels l = 0..3 with one primitive each
    assert list(m.shellsECP) == [4, 4] and list(m.meta["nCore"]) == [2, 2]
    assert list(m.lECP) == [0, 1, 2, 3] * 2 and list(m.KECP) == [1] * 8
    assert list(m.dECP[:4]) == [37.4565, -2.6739, -0.5945, 0.0]
    assert list(m.aECP[:4]) == [6.8446, 7.9317, 6.0209, 1.0]
    assert list(m.nECP) == [2.0] * 8
    # test_c.bs: per atom "6 3", shells s, s, p of three primitives
    assert list(m.shellsBS) == [3, 3] and list(m.lBS) == [0, 0, 1] * 2 and list(m.KBS) == [3] * 6
    assert m.aBS[0] == 153.17226 and m.dBS[0] == 0.07074
    assert m.aBS[8] == 0.232685 and m.dBS[8] == 0.627514
    assert m.nao == 2 * (1 + 1 + 3)


def test_four_column_ecp_records_load_identically(xyz, tmp_path):
    m3 = native.load_molecule(xyz, ECP, BS)
    p = str(tmp_path / "four.ecp")
    with open(p, "w") as f:
        k = 0
        for at in range(2):
            f.write(f"6 2 {m3.shellsECP[at]}\n")
            for s in range(m3.shellsECP[at]):
                sh = at * 4 + s
                f.write(f"  {m3.lECP[sh]} {m3.KECP[sh]}\n")
                for i in range(m3.KECP[sh]):
                    f.write(f"  {i + 1} {m3.aECP[k]:.17g} {m3.dECP[k]:.17g} {m3.nECP[k]:.17g}\n")
                    k += 1
    m4 = native.load_molecule(xyz, p, BS)
    for f in ("shellsECP", "lECP", "KECP", "aECP", "dECP", "nECP"):
        assert np.array_equal(getattr(m3, f), getattr(m4, f)), f


def test_atom_without_ecp_and_blank_lines(tmp_path):
    p = str(tmp_path / "x.xyz")
    write_xyz(p, [0, 0, 0, 1.0, 0, 0, 0, 1.5, 0], comment="")
    e = str(tmp_path / "x.ecp")
    with open(e, "w") as f:
        f.write("6 2 2\n 0 1\n 3.0 1.5 2\n\n 1 2\n 1 0.8 -1.0 2\n 2 0.3 0.25 1\n8 0 0\n1 0 0\n")
    b = str(tmp_path / "x.bs")
    with open(b, "w") as f:
        f.write("6 1\n 0 1\n 1 0.5 1.0\n8 0\n1 2\n 0 1\n 1 1.2 0.7\n 1 2\n 1 0.9 0.3\n 2 0.2 0.6\n")
    m = native.load_molecule(p, e, b)
    assert list(m.shellsECP) == [2, 0, 0] and list(m.lECP) == [0, 1] and list(m.KECP) == [1, 2]
    assert list(m.dECP) == [3.0, -1.0, 0.25] and list(m.aECP) == [1.5, 0.8, 0.3]
    assert list(m.nECP) == [2.0, 2.0, 1.0]
    assert list(m.shellsBS) == [1, 0, 2] and list(m.lBS) == [0, 0, 1] and list(m.KBS) == [1, 1, 2]


@pytest.mark.parametrize("which,text,where", [
    ("xyz", "3\ncomment\nC 0 0 0\nC 1 1 1\n", ":4: fewer atoms"),
    ("xyz", "2\ncomment\nC 0 0 zero\nC 1 1 1\n", ":3: expected 'symbol x y z'"),
    ("ecp", "6 2 1\n 0 1\n 1.0 2.0\n", ":3: expected 'index exponent coefficient power'"),
    ("ecp", "6 2 1\n 0 x\n", ":2: expected 'l K'"),
    ("bs", "6 1\n 0 2\n 1 0.5 1.0\n", "expected 'index exponent coefficient'"),
    ("bs", "6 1\n 0 1\n 1 0.5 1.0\n", "fewer atoms than the geometry"),
])
def test_malformed_files_name_file_and_line(tmp_path, which, text, where):
    p = str(tmp_path / ("bad." + which))
    with open(p, "w") as f:
        f.write(text)
    good_xyz = str(tmp_path / "g.xyz")
    write_xyz(good_xyz, [0, 0, 0, 0, 0, 2.0])
    args = {"xyz": (p, ECP, BS), "ecp": (good_xyz, p, BS), "bs": (good_xyz, ECP, p)}[which]
    with pytest.raises(RuntimeError) as e:
        native.load_molecule(*args)
    assert p in str(e.value) and where in str(e.value), str(e.value)


def test_missing_file():
    with pytest.raises(RuntimeError, match="cannot open"):
        native.load_molecule("/nonexistent/x.xyz", ECP, BS)


def test_loaded_inputs_screen_on_the_host(xyz):
    """The loaded arrays drive libECP_init's host-side screening unchanged."""
    m = native.load_molecule(xyz, ECP, BS)
    h = native.Handle(m)
    try:
        n = native.lib().libecp_screened_triplets(native.C.c_void_p(h.h), None, 0)
    finally:
        h.close()
    assert 0 < n <= m.nominal_triplets()


# ---------------------------------------------------------------- device

def _run_cli(*args):
    r = subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.gpu
def test_cli_matrix_equals_getintegrals(xyz):
    m = native.load_molecule(xyz, ECP, BS)
    rc, I = native.get_integrals(m)
    assert rc == 0
    out = _run_cli(xyz, ECP, BS).splitlines()
    assert out[0].startswith("ECP matrix elements")
    vals = []
    for line in out[1:]:
        if not line.startswith("shell block"):
            vals.extend(float(v) for v in line.split())
    # shell pairs s <= t in AO order, rows of IJK(l_s) x IJK(l_t)
    ls = list(m.lBS)
    off = np.concatenate([[0], np.cumsum([(l + 1) * (l + 2) // 2 for l in ls])])
    ref = []
    for s in range(len(ls)):
        for t in range(s, len(ls)):
            for a in range(off[s], off[s + 1]):
                for b in range(off[t], off[t + 1]):
                    ref.append(I[a, b] if a <= b else 0.0)
    assert np.allclose(vals, ref, rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
def test_cli_energy_and_gradient(xyz, tmp_path):
    m = native.load_molecule(xyz, ECP, BS)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(m.nao, m.nao))
    P = X @ X.T / m.nao
    pf = str(tmp_path / "P.txt")
    np.savetxt(pf, P, fmt="%.17g")
    e = float(_run_cli(xyz, ECP, BS, "-e", pf).split()[1])
    g = np.array([[float(v) for v in line.split()[2:]] for line in _run_cli(xyz, ECP, BS, "-g", pf).splitlines()])
    for n, want in ((0, e), (1, g)):
        h = native.Handle(m, n=n)
        try:
            rc, v = h.contract(P)
        finally:
            h.close()
        assert rc == 0
        assert np.allclose(v, want, rtol=1e-14, atol=0.0)
