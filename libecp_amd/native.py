"""ctypes bindings of the product library ``libecp_amd.so``.

The library exports the reference's C API (``include/libecp.h``,
``include/getIntegrals.h``).  Loading fails loudly when the library has not
been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libecp_amd.so")

CB = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                 C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_void_p)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m libecp_amd.build` first")
        _lib = C.CDLL(LIB_PATH)
        _lib.libECP_init.restype = C.c_void_p
        _lib.libecp_last_error.restype = C.c_char_p
        _lib.libecp_screened_triplets.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        _lib.libecp_last_call_ms.argtypes = [C.c_void_p, C.c_int]
        _lib.libecp_cache_enable.argtypes = [C.c_int]
    return _lib


PHASES = ("tables", "screening", "context", "upload", "kernels", "download", "deliver")


def last_call_ms() -> dict:
    """Phases (ms) of this thread's last getIntegrals / calculateECPIntegrals."""
    out = np.zeros(len(PHASES))
    lib().libecp_last_call_ms(out.ctypes.data, len(PHASES))
    return dict(zip(PHASES, out.tolist()))


def cache_enable(on: bool) -> bool:
    """Enables or disables (and clears) the process cache of tables and
    device contexts; returns the previous setting."""
    return bool(lib().libecp_cache_enable(1 if on else 0))


def cache_clear() -> None:
    lib().libecp_cache_clear()


def last_error() -> str:
    return lib().libecp_last_error().decode()


def _P(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _ijk(l):
    return (l + 1) * (l + 2) // 2


def get_integrals(m, grid=1024, tol=1e-12, acc=1e-14, I=None):
    """getIntegrals() on a :class:`~libecp_amd.workloads.Molecule`; returns
    (rc, matrix) with the upper triangle filled."""
    L = lib()
    nao = m.nao
    if I is None:
        I = np.zeros((nao, nao))
    rc = L.getIntegrals(C.c_int(m.nat), _P(m.geometry, C.c_double), _P(m.shellsECP, C.c_int),
                        _P(m.KECP, C.c_int), _P(m.lECP, C.c_int), _P(m.nECP, C.c_double),
                        _P(m.dECP, C.c_double), _P(m.aECP, C.c_double), _P(m.shellsBS, C.c_int),
                        _P(m.lBS, C.c_int), _P(m.KBS, C.c_int), _P(m.dBS, C.c_double),
                        _P(m.aBS, C.c_double), C.c_int(grid), C.c_double(tol), C.c_double(acc),
                        C.c_int(nao), _P(I, C.c_double))
    return rc, I


def _take(ptr, n, ctype, dtype):
    """Copies n values from a malloc'ed C array and frees it."""
    a = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].astype(dtype, copy=True)
    _libc().free(ptr)
    return a


_libc_h = None


def _libc():
    global _libc_h
    if _libc_h is None:
        _libc_h = C.CDLL(None)
        _libc_h.free.argtypes = [C.c_void_p]
    return _libc_h


def load_molecule(xyz, ecp, basis, name=""):
    """libecp_load_xyz / libecp_load_ecp / libecp_load_basis: the three input
    files of the reference's example program as a Molecule."""
    from .workloads import Molecule
    L = lib()
    nat = C.c_int()
    g = C.POINTER(C.c_double)()
    if L.libecp_load_xyz(xyz.encode(), C.byref(nat), C.byref(g)) < 0:
        raise RuntimeError(last_error())
    n = nat.value
    geometry = _take(g, 3 * n, C.c_double, np.float64)
    ip = [C.POINTER(C.c_int)() for _ in range(4)]
    dp = [C.POINTER(C.c_double)() for _ in range(3)]
    ne = L.libecp_load_ecp(ecp.encode(), nat, C.byref(ip[0]), C.byref(ip[1]), C.byref(ip[2]), C.byref(dp[0]),
                           C.byref(dp[1]), C.byref(dp[2]), C.byref(ip[3]))
    if ne < 0:
        raise RuntimeError(last_error())
    shellsECP = _take(ip[0], n, C.c_int, np.int32)
    lECP = _take(ip[1], ne, C.c_int, np.int32)
    KECP = _take(ip[2], ne, C.c_int, np.int32)
    npe = int(KECP.sum())
    aECP, dECP, nECP = (_take(p, npe, C.c_double, np.float64) for p in dp)
    nCore = _take(ip[3], n, C.c_int, np.int32)
    ib = [C.POINTER(C.c_int)() for _ in range(3)]
    db = [C.POINTER(C.c_double)() for _ in range(2)]
    nb = L.libecp_load_basis(basis.encode(), nat, C.byref(ib[0]), C.byref(ib[1]), C.byref(ib[2]),
                             C.byref(db[0]), C.byref(db[1]))
    if nb < 0:
        raise RuntimeError(last_error())
    shellsBS = _take(ib[0], n, C.c_int, np.int32)
    lBS = _take(ib[1], nb, C.c_int, np.int32)
    KBS = _take(ib[2], nb, C.c_int, np.int32)
    npb = int(KBS.sum())
    aBS, dBS = (_take(p, npb, C.c_double, np.float64) for p in db)
    return Molecule(geometry=geometry, shellsECP=shellsECP, lECP=lECP, KECP=KECP, nECP=nECP, dECP=dECP,
                    aECP=aECP, shellsBS=shellsBS, lBS=lBS, KBS=KBS, dBS=dBS, aBS=aBS, name=name,
                    meta={"nCore": nCore})


class Handle:
    """libECP_init / calculateECPIntegrals / libECP_free."""

    def __init__(self, m, n=0, lmax=-1, ordering=None, grid=1024, tol=1e-12, acc=1e-14):
        L = lib()
        self._m = m  # keep the arrays alive
        self._ord = None if ordering is None else np.ascontiguousarray(ordering, dtype=np.int32)
        op = None if self._ord is None else _P(self._ord, C.c_int)
        self.h = L.libECP_init(C.c_int(m.nat), _P(m.geometry, C.c_double), _P(m.shellsECP, C.c_int),
                               _P(m.lECP, C.c_int), _P(m.KECP, C.c_int), _P(m.nECP, C.c_double),
                               _P(m.dECP, C.c_double), _P(m.aECP, C.c_double),
                               _P(m.shellsBS, C.c_int), _P(m.lBS, C.c_int), _P(m.KBS, C.c_int),
                               _P(m.dBS, C.c_double), _P(m.aBS, C.c_double), C.c_int(n),
                               C.c_int(lmax), op, C.c_int(grid), C.c_double(tol), C.c_double(acc))
        if not self.h:
            raise RuntimeError("libECP_init failed: " + last_error())
        self.n = n

    def blocks(self):
        """Runs calculateECPIntegrals and returns (rc, keys, blocks)."""
        keys, blocks = [], []
        order = self.n

        def cb(A, sa, la, shifta, B, sb, lb, shiftb, Cc, I, args):
            # IJK(la+shifta) x IJK(lb+shiftb); for second derivatives the
            # (1, 0) and (0, 1) blocks keep the unshifted l (libecp.c:364-369)
            las = la if (order == 2 and (shifta, shiftb) == (1, 0)) else la + shifta
            lbs = lb if (order == 2 and (shifta, shiftb) == (0, 1)) else lb + shiftb
            n = _ijk(las) * _ijk(lbs)
            keys.append((A, sa, la, shifta, B, sb, lb, shiftb, Cc))
            blocks.append(np.ctypeslib.as_array(I, shape=(n,)).copy())
        f = CB(cb)
        rc = lib().calculateECPIntegrals(C.c_void_p(self.h), f, None)
        return rc, keys, blocks

    def contract(self, P):
        """libecp_contract: the energy (n = 0, a float) or the gradient
        (n = 1, an array of shape (nat, 3)) of the density P on the device.
        Returns (rc, value)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        out = np.zeros(1 if self.n == 0 else 3 * self._m.nat)
        rc = lib().libecp_contract(C.c_void_p(self.h), C.c_int(P.shape[1]), _P(P, C.c_double),
                                   _P(out, C.c_double))
        if rc < 0:
            raise RuntimeError("libecp_contract failed: " + last_error())
        return rc, (float(out[0]) if self.n == 0 else out.reshape(-1, 3))

    def screened_triplets(self):
        """(C, A, sa, B, sb) of every triplet calculateECPIntegrals evaluates,
        in callback order (host-side screening, no device work)."""
        L = lib()
        n = L.libecp_screened_triplets(self.h, None, 0)
        if n < 0:
            raise RuntimeError(last_error())
        keys = np.zeros((max(n, 1), 5), dtype=np.int32)
        L.libecp_screened_triplets(self.h, keys.ctypes.data, n)
        return keys[:n]

    def close(self):
        if self.h:
            lib().libECP_free(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Mol(C.Structure):
    _fields_ = [("nrAtoms", C.c_int), ("geometry", C.POINTER(C.c_double)),
                ("shellsECP", C.POINTER(C.c_int)), ("KECP", C.POINTER(C.c_int)), ("lECP", C.POINTER(C.c_int)),
                ("nECP", C.POINTER(C.c_double)), ("dECP", C.POINTER(C.c_double)), ("aECP", C.POINTER(C.c_double)),
                ("shellsBS", C.POINTER(C.c_int)), ("lBS", C.POINTER(C.c_int)), ("KBS", C.POINTER(C.c_int)),
                ("dBS", C.POINTER(C.c_double)), ("aBS", C.POINTER(C.c_double))]


def _mol_struct(m):
    return _Mol(m.nat, _P(m.geometry, C.c_double), _P(m.shellsECP, C.c_int), _P(m.KECP, C.c_int),
                _P(m.lECP, C.c_int), _P(m.nECP, C.c_double), _P(m.dECP, C.c_double), _P(m.aECP, C.c_double),
                _P(m.shellsBS, C.c_int), _P(m.lBS, C.c_int), _P(m.KBS, C.c_int), _P(m.dBS, C.c_double),
                _P(m.aBS, C.c_double))


class Batch:
    """libecp_batch_*: many molecules in one resident device job."""

    def __init__(self, mols, grid=1024, tol=1e-12, acc=1e-14, n=0):
        L = lib()
        L.libecp_batch_create_n.restype = C.c_void_p
        L.libecp_batch_triplets.restype = C.c_longlong
        L.libecp_batch_contract.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        self.mols = list(mols)
        self.n = n
        arr = (_Mol * len(self.mols))(*[_mol_struct(m) for m in self.mols])
        self.b = L.libecp_batch_create_n(arr, C.c_int(len(self.mols)), C.c_int(n), C.c_int(grid),
                                         C.c_double(tol), C.c_double(acc))
        if not self.b:
            raise RuntimeError("libecp_batch_create failed: " + last_error())

    def contract_block_doubles(self):
        lib().libecp_batch_block_doubles.restype = C.c_size_t
        return int(lib().libecp_batch_block_doubles(C.c_void_p(self.b)))

    def contract_sizes(self):
        """(doubles of the packed densities, doubles of the packed output)."""
        nP = sum(m.nao * m.nao for m in self.mols)
        nout = len(self.mols) if self.n == 0 else sum(3 * m.nat for m in self.mols)
        return nP, nout

    def contract(self, Ps):
        """libecp_batch_contract with host arrays: Ps is one density per
        molecule.  Returns (codes, outputs, ms): energies (n = 0) or (nat, 3)
        gradients (n = 1) per molecule."""
        P = np.concatenate([np.ascontiguousarray(p, dtype=np.float64).ravel() for p in Ps])
        nP, nout = self.contract_sizes()
        assert P.size == nP
        out = np.zeros(nout)
        codes = np.zeros(len(self.mols), dtype=np.int32)
        ms = C.c_double()
        if lib().libecp_batch_contract(C.c_void_p(self.b), P.ctypes.data, out.ctypes.data, codes.ctypes.data,
                                       0, C.addressof(ms)) != 0:
            raise RuntimeError("libecp_batch_contract failed: " + last_error())
        return codes, self._split(out), ms.value

    def contract_device(self, P_ptr, out_ptr):
        """libecp_batch_contract with device pointers (packed densities in,
        packed results out).  Returns (codes, ms)."""
        codes = np.zeros(len(self.mols), dtype=np.int32)
        ms = C.c_double()
        if lib().libecp_batch_contract(C.c_void_p(self.b), C.c_void_p(P_ptr), C.c_void_p(out_ptr),
                                       codes.ctypes.data, 1, C.addressof(ms)) != 0:
            raise RuntimeError("libecp_batch_contract failed: " + last_error())
        return codes, ms.value

    def _split(self, out):
        if self.n == 0:
            return [float(v) for v in out]
        res, o = [], 0
        for m in self.mols:
            res.append(out[o:o + 3 * m.nat].reshape(-1, 3).copy())
            o += 3 * m.nat
        return res

    def execute(self, repeats=1):
        ms = C.c_double()
        if lib().libecp_batch_execute(C.c_void_p(self.b), C.c_int(repeats), C.byref(ms)) != 0:
            raise RuntimeError("libecp_batch_execute failed: " + last_error())
        return ms.value

    def fetch(self):
        if lib().libecp_batch_fetch(C.c_void_p(self.b)) != 0:
            raise RuntimeError("libecp_batch_fetch failed: " + last_error())

    def matrix(self, i, I=None):
        m = self.mols[i]
        if I is None:
            I = np.zeros((m.nao, m.nao))
        rc = lib().libecp_batch_matrix(C.c_void_p(self.b), C.c_int(i), C.c_int(m.nao), _P(I, C.c_double))
        return rc, I

    def quadrature_ms(self):
        """libecp_batch_quadrature_ms: the type-1 small-grid quadrature's
        device time in the last evaluation (None with per-class launches)."""
        ms, n = C.c_double(), C.c_int()
        if lib().libecp_batch_quadrature_ms(C.c_void_p(self.b), C.byref(ms), C.byref(n)) != 0:
            return None
        self.quadrature_samples = n.value
        return ms.value

    def triplets(self):
        ev = C.c_longlong()
        nom = lib().libecp_batch_triplets(C.c_void_p(self.b), C.byref(ev))
        return nom, ev.value

    def stage_ms(self):
        out = (C.c_double * 8)()
        n = lib().libecp_batch_stage_ms(C.c_void_p(self.b), out, C.c_int(8))
        return list(out[:n])

    def profile(self):
        """{family: (launches, ms, algorithmic_flops)} for one evaluation."""
        L = lib()
        L.libecp_kernel_family.restype = C.c_char_p
        la = (C.c_int * 16)()
        ms = (C.c_double * 16)()
        fl = (C.c_double * 16)()
        k = L.libecp_batch_profile(C.c_void_p(self.b), la, ms, fl, C.c_int(16))
        if k < 0:
            raise RuntimeError("libecp_batch_profile failed: " + last_error())
        return {L.libecp_kernel_family(i).decode(): (la[i], ms[i], fl[i]) for i in range(k)}

    def set_param(self, name, value):
        if lib().libecp_batch_set_param(C.c_void_p(self.b), name.encode(), C.c_int(value)) != 0:
            raise RuntimeError("libecp_batch_set_param failed: " + last_error())

    def get_param(self, name):
        v = C.c_int()
        if lib().libecp_batch_get_param(C.c_void_p(self.b), name.encode(), C.byref(v)) != 0:
            raise RuntimeError("libecp_batch_get_param failed: " + last_error())
        return v.value

    def autotune(self, repeats=3):
        buf = C.create_string_buffer(4096)
        if lib().libecp_batch_autotune(C.c_void_p(self.b), C.c_int(repeats), buf, C.c_int(4096)) != 0:
            raise RuntimeError("libecp_batch_autotune failed: " + last_error())
        return buf.value.decode()

    def close(self):
        if self.b:
            lib().libecp_batch_free(C.c_void_p(self.b))
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """libecp_server_*: requests evaluated with continuous batching on a
    shared device-memory budget.  ``submit`` returns a ticket; ``next``
    returns tickets in completion order; ``result`` collects one."""

    STATS = ("submitted", "finished", "jobs", "requests_in_jobs", "max_jobs_on_device", "max_job_requests",
             "pool_peak_bytes", "pool_bytes", "queued", "running")

    def __init__(self, slots=4, hbm_budget=8 << 30, max_batch_triplets=0, grid=1024, tol=1e-12, acc=1e-14):
        L = lib()
        L.libecp_server_create.restype = C.c_void_p
        L.libecp_server_create.argtypes = [C.c_int, C.c_size_t, C.c_longlong, C.c_int, C.c_double, C.c_double]
        L.libecp_server_submit.restype = C.c_longlong
        L.libecp_server_submit.argtypes = [C.c_void_p, C.POINTER(_Mol)]
        L.libecp_server_wait.argtypes = [C.c_void_p, C.c_longlong, C.c_double]
        L.libecp_server_next.restype = C.c_longlong
        L.libecp_server_next.argtypes = [C.c_void_p, C.c_double]
        L.libecp_server_result.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_void_p, C.c_void_p]
        L.libecp_server_stats.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.libecp_server_free.argtypes = [C.c_void_p]
        self.s = L.libecp_server_create(slots, hbm_budget, max_batch_triplets, grid, tol, acc)
        if not self.s:
            raise RuntimeError("libecp_server_create failed: " + last_error())
        self.nao = {}

    def submit(self, m):
        st = _mol_struct(m)
        t = lib().libecp_server_submit(self.s, C.byref(st))
        if t < 0:
            raise RuntimeError("libecp_server_submit failed: " + last_error())
        self.nao[t] = m.nao
        return t

    def wait(self, ticket, timeout=-1.0):
        return lib().libecp_server_wait(self.s, ticket, timeout)

    def next(self, timeout=-1.0):
        t = lib().libecp_server_next(self.s, timeout)
        return None if t < 0 else t

    def result(self, ticket, matrix=True):
        """(rc, matrix or None, (t_submit, t_admit, t_done)); rc < 0 carries
        the error text as a fourth element."""
        nao = self.nao.pop(ticket)
        I = np.zeros((nao, nao)) if matrix else None
        times = (C.c_double * 3)()
        rc = lib().libecp_server_result(self.s, ticket, nao, None if I is None else I.ctypes.data, times)
        out = (rc, I, tuple(times))
        return out + (last_error(),) if rc < 0 else out

    def stats(self):
        v = (C.c_longlong * 16)()
        k = lib().libecp_server_stats(self.s, v, 16)
        return dict(zip(self.STATS, v[:k]))

    def close(self):
        if self.s:
            lib().libecp_server_free(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
