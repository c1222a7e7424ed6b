"""libecp_batch_contract_noalloc: the density contraction with every device
buffer beyond the inputs and outputs in the library's static scratch array.

Checked against libecp_batch_contract (device buffers) on two calls in a
row, the second with a smaller input (a batch of fewer molecules), for the
energy (n = 0) and the gradient (n = 1): bit for bit, since both run the
same kernels over the same blocks in the same order; and the regular entry
point after the no-allocation one on the same batch.  (The rejection of a
job larger than the scratch needs millions of triplets and is not run.)
"""
import ctypes as C

import numpy as np
import pytest

from libecp_amd import native
from libecp_amd.workloads import medium_cluster, uneven_batch

pytestmark = pytest.mark.gpu


class Dev:
    """Device buffers of the test itself (the caller's inputs and outputs)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(8, a.nbytes))) == 0
        if a.nbytes:
            assert self.hip.hipMemcpy(p, C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1) == 0
        self.bufs.append(p)
        return p.value

    def get(self, ptr, n):
        out = np.zeros(n)
        assert self.hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(8 * n), 2) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)


def densities(mols, seed):
    rng = np.random.default_rng(seed)
    Ps = []
    for m in mols:
        X = rng.normal(size=(m.nao, max(1, m.nao // 2)))
        Ps.append(X @ X.T / m.nao)
    return np.concatenate([P.reshape(-1) for P in Ps])


def noalloc(b, P_ptr, out_ptr, nmol):
    L = native.lib()
    L.libecp_batch_contract_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    codes = np.full(nmol, -7, dtype=np.int32)
    ms = C.c_double()
    rc = L.libecp_batch_contract_noalloc(C.c_void_p(b.b), C.c_void_p(P_ptr), C.c_void_p(out_ptr),
                                         C.c_void_p(codes.ctypes.data), C.byref(ms))
    return rc, codes


@pytest.mark.parametrize("n", [0, 1])
def test_noalloc_equals_contract_twice_second_smaller(n):
    big = uneven_batch(0)[:24]
    small = big[:5]
    dev = Dev()
    try:
        for mols, seed in ((big, 1), (small, 2)):
            b = native.Batch(mols, n=n)
            try:
                b.execute(1)
                nP, nout = b.contract_sizes()
                P = densities(mols, seed)
                assert P.size == nP
                Pd = dev.put(P)
                o1 = dev.put(np.full(nout, np.nan))
                o2 = dev.put(np.full(nout, np.nan))
                b.contract_device(Pd, o1)
                rc, codes = noalloc(b, Pd, o2, len(mols))
                assert rc == 0, native.last_error()
                assert np.all(codes == 0)
                r1, r2 = dev.get(o1, nout), dev.get(o2, nout)
                assert np.all(np.isfinite(r2))
                assert np.array_equal(r1, r2), float(np.abs(r1 - r2).max())
            finally:
                b.close()
    finally:
        dev.close()


def test_noalloc_then_regular_entry_point():
    m = medium_cluster(0)
    dev = Dev()
    try:
        b = native.Batch([m], n=0)
        try:
            b.execute(1)
            nP, nout = b.contract_sizes()
            P = densities([m], 3)
            Pd = dev.put(P)
            o2 = dev.put(np.zeros(nout))
            rc, codes = noalloc(b, Pd, o2, 1)
            assert rc == 0, native.last_error()
            codes_r, outs, _ = b.contract([P.reshape(m.nao, m.nao)])
            assert codes_r[0] == codes[0] == 0
            assert outs[0] == dev.get(o2, nout)[0]
        finally:
            b.close()
    finally:
        dev.close()
