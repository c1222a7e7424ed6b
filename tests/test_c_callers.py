"""C programs that call getIntegrals() like the reference's example
(example/ex1.c:162), linked against libecp_amd.so and against the static
archive libecp_amd.a (tests/c/example_getintegrals.c, built by
libecp_amd/build.py)."""
import os
import subprocess

import numpy as np
import pytest

import ecp_native as E
from conftest import within
from libecp_amd import native
from libecp_amd.build import EXAMPLES
from libecp_amd.workloads import plain_small


def molecule_text(m):
    nprimE = len(m.nECP)
    parts = [m.nat, *m.geometry, *m.shellsECP, len(m.KECP), *m.KECP, *m.lECP, nprimE, *m.nECP, *m.dECP,
             *m.aECP, *m.shellsBS, len(m.lBS), *m.lBS, *m.KBS, len(m.dBS), *m.dBS, *m.aBS, m.nao]
    return " ".join(repr(float(v)) if isinstance(v, (float, np.floating)) else str(int(v)) for v in parts)


def run_example(kind, m):
    p = subprocess.run([EXAMPLES[kind]], input=molecule_text(m), capture_output=True, text=True, timeout=120)
    lines = p.stdout.split()
    rc = int(lines[0]) if lines else None
    vals = np.array([float(v) for v in lines[1:]])
    return p.returncode, rc, vals, p.stderr


def gpu_present():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.parametrize("kind", ["shared", "static"])
def test_examples_are_built(kind):
    assert os.access(EXAMPLES[kind], os.X_OK), EXAMPLES[kind]


@pytest.mark.skipif(gpu_present(), reason="checks the behaviour without a GPU")
@pytest.mark.parametrize("kind", ["shared", "static"])
def test_examples_fail_loudly_without_device(kind):
    code, rc, vals, err = run_example(kind, plain_small(0))
    assert rc == 1 and code == 1
    assert "libecp:" in err


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["shared", "static"])
def test_examples_match_library_and_oracle(kind):
    m = plain_small(0)
    code, rc, vals, err = run_example(kind, m)
    assert code == 0 and rc == 0, err
    rc1, I = native.get_integrals(m)
    up = I[np.triu_indices(m.nao)]
    assert np.array_equal(vals, up)
    rc2, Io = E.oracle_getIntegrals(m, options=1)
    assert within(vals, Io[np.triu_indices(m.nao)], 1e-10, 1e-8)[0]
