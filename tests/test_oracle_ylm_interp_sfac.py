"""Pins the oracle's real Y_lm (test/math/test_sph_harmonics.cpp), beta(q)
interpolation (src/pseudo/ncpp_advance.cppm:644-674) and structure factor
(src/H_psi/structure_factor.cppm:92-159)."""
import math

import numpy as np
import pytest

import oracle
import refmath
from effmass_amd import synth


def test_ylm_kat():                                      # test_sph_harmonics.cpp:20-60
    y, _, _ = oracle.ylm(0.0, 0.0, 3)
    assert abs(y[0] - 1 / (2 * math.sqrt(math.pi))) < 1e-12
    y, _, _ = oracle.ylm(0.5, 0.0, 3)
    assert abs(y[2] - math.sqrt(3 / (4 * math.pi)) * math.cos(0.5)) < 1e-12   # Y10
    y, _, _ = oracle.ylm(0.5, 0.3, 3)
    c = -math.sqrt(3 / (4 * math.pi)) * math.sin(0.5)
    assert abs(y[3] - c * math.cos(0.3)) < 1e-12        # Y11 (CS phase)
    assert abs(y[1] - c * math.sin(0.3)) < 1e-12        # Y1-1


def test_ylm_theta0():                                   # test_sph_harmonics.cpp:135-139
    y, _, _ = oracle.ylm(0.0, 0.3, 5)
    assert abs(y[25 + 5] - math.sqrt(11 / (4 * math.pi))) < 1e-12
    assert abs(y[25 + 5 + 3]) < 1e-12


@pytest.mark.parametrize("theta,phi", [(0.7, 0.3), (2.1, -2.5), (1e-3, 1.0), (math.pi / 2, 0.0)])
def test_ylm_vs_scipy(theta, phi):
    L = 5
    y, _, _ = oracle.ylm(theta, phi, L)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            ref = refmath.real_ylm(l, m, np.array([theta]), np.array([phi]))[0]
            assert abs(y[l * l + m + l] - ref) < 1e-13, (l, m)


def test_ylm_gradients_fd():                             # test_sph_harmonics.cpp:202-221, 379
    th, ph, L, h = 0.7, 0.3, 4, 1e-6
    _, dt, dp = oracle.ylm(th, ph, L)
    yp = oracle.ylm(th + h, ph, L)[0]; ym = oracle.ylm(th - h, ph, L)[0]
    assert np.allclose((yp - ym) / (2 * h), dt, atol=1e-8)
    yp = oracle.ylm(th, ph + h, L)[0]; ym = oracle.ylm(th, ph - h, L)[0]
    assert np.allclose((yp - ym) / (2 * h) / math.sin(th), dp, atol=1e-8)


def test_ylm_grad_theta0_limits():                       # test_sph_harmonics.cpp:341-343, 403-412
    _, dt, dp = oracle.ylm(0.0, 0.3, 4)
    c = -math.sqrt(3 / (4 * math.pi))
    assert abs(dt[3] - c * math.cos(0.3)) < 1e-12 and abs(dt[1] - c * math.sin(0.3)) < 1e-12
    for l in range(5):
        assert abs(dt[l * l + l]) < 1e-12


def _species():
    return synth.make_species(synth.SplitMix64(7))


def test_interp_nodes_and_rules():
    sp = _species()
    t = sp.beta_tab[2]
    norm = 0.37
    for i in (1, 2, 17, 500, 1999):
        rc, v = oracle.interp(t, sp.dq, sp.qmax, norm, i * sp.dq)
        assert rc == 0 and abs(v - t[i] * norm) <= 1e-13 * (abs(t).max() * norm)
    rc, v = oracle.interp(t, sp.dq, sp.qmax, norm, 0.0)
    assert rc == 0 and v == t[0] * norm                 # q == 0 returns node 0
    assert oracle.interp(t, sp.dq, sp.qmax, norm, -1e-3)[0] != 0      # domain_error
    assert oracle.interp(t, sp.dq, sp.qmax, norm, sp.qmax + 1e-9)[0] != 0


def test_interp_vs_closed_form():
    sp = _species()
    q = np.linspace(0.0, 12.0, 997)[1:]
    for k in range(sp.nbeta):
        ex = sp.beta_exact(k, q)
        got = np.array([oracle.interp(sp.beta_tab[k], sp.dq, sp.qmax, 1.0, x)[1] for x in q])
        assert np.abs(got - ex).max() < 1e-8 * np.abs(ex).max()
        exd = sp.beta_exact(k, q, deriv=True)
        gotd = np.array([oracle.interp(sp.dbeta_tab[k], sp.dq, sp.qmax, 1.0, x)[1] for x in q])
        assert np.abs(gotd - exd).max() < 1e-8 * np.abs(exd).max()


def test_sfac():
    n = (36, 36, 36)
    assert abs(oracle.sfac([0, 0, 0], [3, -4, 7], [0, 0, 0], n) - 1) < 1e-15
    rng = np.random.default_rng(1)
    for _ in range(20):
        tau = rng.uniform(-0.5, 0.5, 3)
        g = rng.integers(-18, 18, 3)
        kf = rng.uniform(-0.5, 0.5, 3)
        s = oracle.sfac(tau, g, kf, n)
        assert abs(abs(s) - 1) < 1e-14
        assert abs(s - np.exp(-2j * np.pi * np.dot(g + kf, tau))) < 1e-13


def test_interp_stencil_centring():
    """Quartic node data tells the stencil apart: a cubic Lagrange
    interpolant through nodes j_0..j_3 misses f(x) = (x - c)^4 (x = q/dq in
    node units) by exactly prod_k (x - j_k), since the fourth derivative is
    the constant 4!.  The reference centres the stencil on
    i0 = floor(q/dq) - 1, clamped to 0 below the second node
    (src/pseudo/ncpp_advance.cppm:663-673); an uncentred stencil
    i0 = floor(q/dq) has another product and fails."""
    dq, n = 0.01, 2004
    qmax = (n - 4) * dq
    for q in (0.0037, 0.0123, 0.0251, 0.5043, 1.234567, 7.77777, 19.5071):
        s = q / dq
        c = math.floor(s)
        t = (np.arange(n, dtype=np.float64) - c) ** 4
        i0 = max(c - 1, 0)
        want = (s - c) ** 4 - np.prod(s - (i0 + np.arange(4)))
        rc, got = oracle.interp(t, dq, qmax, 1.0, q)
        assert rc == 0
        assert abs(got - want) <= 1e-9, q
        if c >= 1:
            # past the first node the uncentred choice misses by O(0.1)
            off = (s - c) ** 4 - np.prod(s - (c + np.arange(4)))
            assert abs(off - want) > 1e-2, q
