"""Request server (libecp_server_*): continuous batching on a shared HBM
budget.  Every request's matrix must equal the one getIntegrals() gives for
the same molecule alone -- bit for bit, since each triplet runs the same
kernels with the same summation order whatever it is batched with -- and
the server must reject a request that cannot fit its budget without
disturbing the others.
"""
import numpy as np
import pytest

from libecp_amd import native
from libecp_amd.workloads import batch_molecule, perturb


def _requests(seed, n, nat_range, frac_perturbed=0.3):
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(n):
        if mols and rng.random() < frac_perturbed:
            mols.append(perturb(rng, mols[int(rng.integers(len(mols)))]))
        else:
            mols.append(batch_molecule(rng, int(rng.integers(nat_range[0], nat_range[1] + 1))))
    return mols


def _collect(S, tickets, timeout=60.0):
    got = {}
    while len(got) < len(tickets):
        t = S.next(timeout)
        assert t is not None, "server produced no result within the time limit"
        got[t] = S.result(t)
    return got


@pytest.mark.gpu
def test_server_mix_equals_alone():
    mols = _requests(seed=7, n=48, nat_range=(3, 30))
    # a small batch cap forces many jobs, so requests overlap on the device
    S = native.Server(slots=3, hbm_budget=1 << 30, max_batch_triplets=40000)
    try:
        tickets = [S.submit(m) for m in mols]
        got = _collect(S, tickets)
        st = S.stats()
    finally:
        S.close()
    assert st["finished"] == len(mols) and st["jobs"] > 1
    assert st["max_jobs_on_device"] >= 2, st
    for t, m in zip(tickets, mols):
        rc, I = native.get_integrals(m)
        rc2, I2, times = got[t][:3]
        assert rc2 == 0 and rc == 0
        assert times[0] <= times[1] <= times[2]
        # same alone and in the mix: tolerance 0 (bitwise)
        assert np.array_equal(I, I2), (m.nat, float(np.abs(I - I2).max()))


@pytest.mark.gpu
def test_server_budget_rejects_oversized_request():
    small = _requests(seed=3, n=6, nat_range=(3, 6), frac_perturbed=0.0)
    big = batch_molecule(np.random.default_rng(11), 60)
    S = native.Server(slots=2, hbm_budget=4 << 20, max_batch_triplets=0)
    try:
        tickets = [S.submit(m) for m in small[:3]] + [S.submit(big)] + [S.submit(m) for m in small[3:]]
        got = _collect(S, tickets)
    finally:
        S.close()
    order = small[:3] + [big] + small[3:]
    for t, m in zip(tickets, order):
        r = got[t]
        if m is big:
            assert r[0] == -1 and "HBM budget" in r[3]
        else:
            rc, I = native.get_integrals(m)
            assert r[0] == 0 and np.array_equal(I, r[1])


@pytest.mark.gpu
def test_server_wait_and_result_errors():
    m = _requests(seed=5, n=1, nat_range=(4, 4))[0]
    S = native.Server(slots=1, hbm_budget=64 << 20)
    try:
        t = S.submit(m)
        assert S.wait(t, 60.0) == 0
        assert S.wait(t + 100, 0.0) == -1
        rc, I, times = S.result(t)
        assert rc == 0
        rc_ref, I_ref = native.get_integrals(m)
        assert np.array_equal(I, I_ref)
        # a collected ticket is gone
        S.nao[t] = m.nao
        assert S.result(t)[0] == -1
    finally:
        S.close()
