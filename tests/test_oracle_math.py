"""Pins the CPU oracle with mathematics that does not depend on the
reference: closed-form grid nodes and weights, Gaussian moments through the
adaptive integrators, scipy's Bessel functions and spherical harmonics,
exact three-centre Gaussian integrals for the local (type-1) channel, exact
one-centre integrals and an independent high-order 3-D quadrature for the
semi-local (type-2) projectors."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
import scipy.special as sp

import ecp_native as E
from cases import _mol

lib = E.load(E.ORACLE_SO)
FN = C.CFUNCTYPE(C.c_double, C.c_double)
lib.orc_quad_ps93.restype = C.c_double
lib.orc_quad_psm92.restype = C.c_double


# ---------------------------------------------------------------- grids ----

def test_ps93_grid_closed_form():
    """383 Gauss-Chebyshev (2nd kind, PSM-transformed) nodes
    x_k = 1 - 2k/384 + (2/pi)(1 + 2/3 sin^2 t) cos t sin t, t = k pi/384,
    weights sin^4 t, then r = 1 - log2(1-x), w' = w / (ln2 (1-x))."""
    x = np.zeros(2047); w = np.zeros(2047); n = C.c_int()
    lib.orc_grid_ps93(128, E._P(x, C.c_double), E._P(w, C.c_double), C.byref(n))
    assert n.value == 383
    k = np.arange(1, 384)
    t = k * np.pi / 384
    xc = 1 - 2 * k / 384 + (2 / np.pi) * (1 + 2 / 3 * np.sin(t) ** 2) * np.cos(t) * np.sin(t)
    wc = np.sin(t) ** 4
    order = np.argsort(xc)
    xc, wc = xc[order], wc[order]
    # compare on (-1, 1): the map's derivative 1/(ln2 (1-x)) amplifies an ulp
    # of x near x = 1 into ~1e-9 in r
    xb = 1 - 2.0 ** (1 - x[:383])
    assert np.abs(xb - xc).max() < 1e-14
    assert np.allclose(w[:383] * np.log(2) * (1 - xb), wc, rtol=1e-12, atol=0)
    assert np.all(np.diff(x[:383]) > 0)


def test_psm92_grid_closed_form():
    x = np.zeros(2047); w = np.zeros(2047); n = C.c_int()
    lib.orc_grid_psm92(1024, E._P(x, C.c_double), E._P(w, C.c_double), C.byref(n))
    assert n.value == 1023
    k = np.arange(1, 1024)
    t = k * np.pi / 1024
    xc = 1 - 2 * k / 1024 + (2 / np.pi) * (1 + 2 / 3 * np.sin(t) ** 2) * np.cos(t) * np.sin(t)
    order = np.argsort(xc)
    assert np.allclose(x[:1023], xc[order], rtol=1e-13, atol=1e-15)
    assert np.allclose(w[:1023], (np.sin(t) ** 4)[order], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("k,c", [(0, 1.0), (2, 1.0), (4, 0.7), (6, 2.5), (3, 0.3)])
def test_ps93_integrates_gaussian_moments(k, c):
    """int_0^inf r^k exp(-c r^2) dr = Gamma((k+1)/2) / (2 c^((k+1)/2))."""
    f = FN(lambda r: r ** k * math.exp(-c * r * r))
    rc = C.c_int()
    val = lib.orc_quad_ps93(f, C.c_double(1e-12), C.byref(rc))
    exact = math.gamma((k + 1) / 2) / (2 * c ** ((k + 1) / 2))
    assert rc.value == 0
    assert abs(val - exact) < 1e-11 * max(1.0, exact)


@pytest.mark.parametrize("zeta,P,k", [(3.0, 2.0, 2), (40.0, 1.2, 4), (0.5, 0.0, 1)])
def test_psm92_fm06_integrates_shifted_gaussian(zeta, P, k):
    """int_0^inf r^k exp(-zeta (r-P)^2) dr on [max(0,P-7s), P+9s], by direct
    high-order Gauss-Legendre as the independent value."""
    f = FN(lambda r: r ** k * math.exp(-zeta * (r - P) ** 2))
    rc = C.c_int()
    val = lib.orc_quad_psm92(1024, C.c_double(zeta), C.c_double(P), f, C.c_double(1e-12), C.byref(rc))
    s = 1 / math.sqrt(zeta)
    lo, hi = max(0.0, P - 7 * s), P + 9 * s
    xg, wg = np.polynomial.legendre.leggauss(400)
    rr = 0.5 * (hi - lo) * xg + 0.5 * (hi + lo)
    exact = 0.5 * (hi - lo) * np.sum(wg * rr ** k * np.exp(-zeta * (rr - P) ** 2))
    assert rc.value == 0
    assert abs(val - exact) < 1e-11 * max(1.0, abs(exact))


# --------------------------------------------------------------- Bessel ----

@pytest.mark.parametrize("z", [0.0, 3e-8, 1e-3, 0.37, 1.0, 2.345, 7.777, 12.5, 15.999, 16.0, 23.4, 120.0])
def test_weighted_bessel_against_scipy(z):
    """K_l(z) = exp(-z) i_l(z); all three branches of bessel.c:102-199."""
    lmax = 12
    K = np.zeros(lmax + 1)
    lib.orc_bessel_eval(C.c_int(20), C.c_double(1e-14), C.c_int(lmax), C.c_double(z), E._P(K, C.c_double))
    ref = np.array([sp.spherical_in(l, z) * math.exp(-z) if z > 0 else float(l == 0) for l in range(lmax + 1)])
    assert np.allclose(K, ref, rtol=1e-10, atol=2e-14), (K - ref)


def test_bessel_table_rows_against_scipy():
    rows, N = 10, 1600
    tab = np.zeros((rows + 1) * (N + 1))
    assert lib.orc_bessel_table(C.c_int(rows), C.c_int(N), C.c_int(200), C.c_double(1e-14), E._P(tab, C.c_double)) == 0
    tab = tab.reshape(rows + 1, N + 1)
    z = np.arange(1, N + 1) / 100.0
    for l in range(rows + 1):
        ref = sp.spherical_in(l, z) * np.exp(-z)
        assert np.allclose(tab[l, 1:], ref, rtol=1e-11, atol=3e-14), l


def test_bessel_accuracy_below_zero_fails_tabulation():
    """accuracy < 0 never ends the series: the cut-off is reached and the
    tabulation reports failure (bessel.c:51-56)."""
    tab = np.zeros(11 * 1601)
    assert lib.orc_bessel_table(C.c_int(10), C.c_int(1600), C.c_int(200), C.c_double(-1.0), E._P(tab, C.c_double)) == 1


# ---------------------------------------------------- spherical harmonics --

def real_sph_scipy(l, m, theta, phi):
    """Real spherical harmonics without the Condon-Shortley phase."""
    y = sp.sph_harm_y(l, abs(m), theta, phi)
    if m == 0:
        return y.real
    s = math.sqrt(2) * (-1) ** m
    return s * (y.real if m > 0 else y.imag)


@pytest.mark.parametrize("theta,phi", [(0.3, 1.1), (2.7, 5.9), (0.0, 0.0), (np.pi / 2, 2.0), (1.9, 0.0)])
def test_real_spherical_harmonics_against_scipy(theta, phi):
    lmax = 9
    out = np.zeros((lmax + 1) ** 2)
    lib.orc_real_sph(C.c_int(lmax), C.c_double(theta), C.c_double(phi), E._P(out, C.c_double))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            assert abs(out[l * l + l + m] - real_sph_scipy(l, m, theta, phi)) < 1e-12, (l, m)


# ------------------------------------------------------ exact integrals ----

def cart_exps(l):
    """libint order of dimensions.c:17-38."""
    return [(l - i, i - j, j) for i in range(l + 1) for j in range(i + 1)]


def gauss_moment(m, p):
    return 0.0 if m % 2 else math.gamma((m + 1) / 2) / p ** ((m + 1) / 2)


def overlap3_1d(i, j, A, B, Cc, al, be, ga):
    """int (x-A)^i (x-B)^j exp(-al(x-A)^2 - be(x-B)^2 - ga(x-C)^2) dx."""
    p = al + be + ga
    P = (al * A + be * B + ga * Cc) / p
    K = math.exp(-(al * be * (A - B) ** 2 + al * ga * (A - Cc) ** 2 + be * ga * (B - Cc) ** 2) / p)
    s = 0.0
    for a in range(i + 1):
        for b in range(j + 1):
            s += math.comb(i, a) * math.comb(j, b) * (P - A) ** (i - a) * (P - B) ** (j - b) * gauss_moment(a + b, p)
    return K * s


def shells_of(m):
    out, k, ps = [], 0, 0
    for at, ns in enumerate(m.shellsBS):
        for _ in range(ns):
            l, K = int(m.lBS[k]), int(m.KBS[k])
            out.append((at, l, m.aBS[ps:ps + K], m.dBS[ps:ps + K]))
            ps += K
            k += 1
    return out


def type_blocks(m, t):
    rc, col = E.oracle_blocks(m)
    assert rc == 0
    return {k: b for i, (k, b) in enumerate(zip(col.keys, col.blocks)) if i % 2 == t - 1}


def test_type1_equals_three_centre_gaussian_integral():
    """With n = 2 the local channel is a Gaussian potential, so the type-1
    block is the exact three-centre overlap of Gaussians (any geometry)."""
    local = [(2, 1.7, 0.8), (2, -0.6, 2.3)]
    ecp = [(0, [(2, 1.0, 1.0)]), (1, [(2, 1.0, 1.0)]), (2, local)]
    basis = [(0, [(1.3, 0.6), (0.4, 0.5)]), (1, [(0.9, 0.7), (0.3, 0.4)]), (2, [(1.1, 0.8)])]
    geom = [0.2, -0.1, 0.3, 1.1, 0.6, -0.4, -0.7, 0.9, 0.8]
    m = _mol(geom, [basis, basis, basis], [ecp, None, None])
    blocks = type_blocks(m, 1)
    sh = shells_of(m)
    G = m.geometry.reshape(-1, 3)
    Cc = G[0]
    worst = 0.0
    for (A, sa, la, _, B, sb, lb, _, _), blk in blocks.items():
        ia = sum(m.shellsBS[:A]) + sa
        ib = sum(m.shellsBS[:B]) + sb
        _, _, aa, da = sh[ia]
        _, _, ab, db = sh[ib]
        exact = np.zeros((len(cart_exps(la)), len(cart_exps(lb))))
        for i, ea in enumerate(cart_exps(la)):
            for j, eb in enumerate(cart_exps(lb)):
                v = 0.0
                for x1, c1 in zip(aa, da):
                    for x2, c2 in zip(ab, db):
                        for (n, d, g) in local:
                            prod = 1.0
                            for ax in range(3):
                                prod *= overlap3_1d(ea[ax], eb[ax], G[A][ax], G[B][ax], Cc[ax], x1, x2, g)
                            v += c1 * c2 * d * prod
                exact[i, j] = v
        err = np.abs(blk.reshape(exact.shape) - exact).max() / max(1e-3, np.abs(exact).max())
        worst = max(worst, err)
    assert worst < 1e-9, worst


def sphere_rule(nt=48, nph=96):
    xg, wg = np.polynomial.legendre.leggauss(nt)
    th = np.arccos(xg)
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    W = np.outer(wg, np.full(nph, 2 * np.pi / nph))
    return T.ravel(), Ph.ravel(), W.ravel()


def real_sph_grid(l, T, Ph):
    return np.array([[real_sph_scipy(l, mm, t, p) for t, p in zip(T, Ph)] for mm in range(-l, l + 1)])


def test_type2_one_centre_closed_form():
    """A = B = C: sum_l sum_m <mono_a|S_lm><S_lm|mono_b> times the radial
    moment d Gamma((la+lb+n+1)/2) / (2 (alpha+beta+a)^((la+lb+n+1)/2))."""
    ecp = [(0, [(2, 2.5, 1.4), (1, -0.7, 0.6)]), (1, [(2, -1.3, 0.9)]), (2, [(0, 0.6, 2.1), (2, 0.4, 0.5)]),
           (3, [(2, 0.2, 1.0)])]
    basis = [(0, [(1.3, 0.6), (0.4, 0.5)]), (1, [(0.9, 0.7)]), (2, [(1.1, 0.8), (0.35, 0.2)])]
    m = _mol([0.0, 0.0, 0.0], [basis], [ecp])
    blocks = type_blocks(m, 2)
    sh = shells_of(m)
    T, Ph, W = sphere_rule()
    X, Y, Z = np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)
    Sl = {l: real_sph_grid(l, T, Ph) for l in range(3)}
    for (A, sa, la, _, B, sb, lb, _, _), blk in blocks.items():
        _, _, aa, da = sh[sa]
        _, _, ab, db = sh[sb]
        ma = np.array([X ** e[0] * Y ** e[1] * Z ** e[2] for e in cart_exps(la)])
        mb = np.array([X ** e[0] * Y ** e[1] * Z ** e[2] for e in cart_exps(lb)])
        exact = np.zeros((ma.shape[0], mb.shape[0]))
        for l, prims in ecp[:3]:
            Pa = (ma * W) @ Sl[l].T
            Pb = (mb * W) @ Sl[l].T
            ang = Pa @ Pb.T
            rad = 0.0
            for x1, c1 in zip(aa, da):
                for x2, c2 in zip(ab, db):
                    for n, d, g in prims:
                        k = la + lb + n
                        rad += c1 * c2 * d * math.gamma((k + 1) / 2) / (2 * (x1 + x2 + g) ** ((k + 1) / 2))
            exact += ang * rad
        err = np.abs(blk.reshape(exact.shape) - exact).max() / max(1e-3, np.abs(exact).max())
        assert err < 1e-9, ((la, lb), err)


def test_type2_general_geometry_against_3d_quadrature():
    """<a| U_l(r) P_l |b> = int_0^inf dr U_l(r) sum_m A_lm(r) B_lm(r) with
    A_lm(r) = int dOmega S_lm chi_a(C + r Omega); Gauss-Legendre in r times a
    product rule on the sphere, independent of the Bessel expansion."""
    ecp = [(0, [(2, 2.0, 1.1)]), (1, [(2, -1.1, 0.7)]), (2, [(2, 0.9, 1.3)])]
    basis = [(0, [(0.8, 0.6), (0.3, 0.5)]), (1, [(0.6, 0.7)])]
    geom = [0.0, 0.0, 0.0, 0.9, -0.4, 0.5, -0.6, 0.8, -0.3]
    m = _mol(geom, [basis, basis, basis], [ecp, None, None])
    blocks = type_blocks(m, 2)
    sh = shells_of(m)
    G = m.geometry.reshape(-1, 3)
    T, Ph, W = sphere_rule(40, 80)
    U = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)], 1)
    xr, wr = np.polynomial.legendre.leggauss(160)
    R = 14.0
    r = 0.5 * R * (xr + 1)
    wrr = 0.5 * R * wr
    Sl = {l: real_sph_grid(l, T, Ph) for l in range(2)}
    pts = G[0][None, None, :] + r[:, None, None] * U[None, :, :]   # (nr, nang, 3)

    def proj(ishell):
        at, l, a, d = sh[ishell]
        rel = pts - G[at]
        r2 = (rel ** 2).sum(-1)
        rad = sum(c * np.exp(-x * r2) for x, c in zip(a, d))
        vals = np.array([rel[..., 0] ** e[0] * rel[..., 1] ** e[1] * rel[..., 2] ** e[2] * rad for e in cart_exps(l)])
        return {ll: np.einsum("cra,ma->crm", vals * W, Sl[ll]) for ll in Sl}

    cache = {}
    worst = 0.0
    for (A, sa, la, _, B, sb, lb, _, _), blk in blocks.items():
        ia = sum(m.shellsBS[:A]) + sa
        ib = sum(m.shellsBS[:B]) + sb
        for i in (ia, ib):
            if i not in cache:
                cache[i] = proj(i)
        exact = 0.0
        for l, prims in ecp[:2]:
            Ul = sum(d * r ** n * np.exp(-g * r * r) for n, d, g in prims)
            exact = exact + np.einsum("r,arm,brm->ab", wrr * Ul, cache[ia][l], cache[ib][l])
        err = np.abs(blk.reshape(exact.shape) - exact).max() / max(1e-3, np.abs(exact).max())
        worst = max(worst, err)
    assert worst < 1e-8, worst
