"""k.p effective mass (effmass_amd/effmass.py) pinned by the mathematics:
with all states of a dense H(k) in the sum, (1/m*) equals the finite-
difference second derivative of the eigenvalue E_n(k) in the same G basis;
free electrons give 1/m* = identity."""
import copy

import numpy as np

import refmath
from effmass_amd import effmass, synth


def _cell():
    return synth.custom(*synth.fcc_primitive(10.26), (24, 24, 24), 12.0,
                        [np.array([0.11, 0.23, -0.17])], 2, seed=21)


def dense_dh(cell, kp):
    P, lam, dP = refmath.projectors(cell, kp, deriv=True)
    return [np.diag(kp.K[:, a]) + (dP[:, a].T * lam) @ P.conj() + (P.T * lam) @ dP[:, a].conj()
            for a in range(3)]


def test_kp_sum_equals_eigenvalue_curvature():
    wl = _cell()
    cell, kp = wl.cell, wl.kpts[0]
    A = cell.A
    E, X = np.linalg.eigh(refmath.dense_h(cell, kp))          # columns = states
    dH = dense_dh(cell, kp)
    V = [X.conj().T @ d @ X for d in dH]                       # <m|dH_a|n>
    delta = 2e-3
    bands = [0, 3]
    d2 = np.zeros((len(bands), 3, 3))
    for b in range(3):
        dk = np.zeros(3); dk[b] = delta
        dp = dense_dh(cell, effmass.shifted_kpoint(kp, A, dk))
        dm = dense_dh(cell, effmass.shifted_kpoint(kp, A, -dk))
        for i, n in enumerate(bands):
            for a in range(3):
                d2[i, a, b] = np.real(X[:, n].conj() @ ((dp[a] - dm[a]) / (2 * delta)) @ X[:, n])
    # finite-difference Hessian of E_n(k) in the same basis
    def En(dk):
        return np.linalg.eigvalsh(refmath.dense_h(cell, effmass.shifted_kpoint(kp, A, dk)))
    h = 2e-3
    for i, n in enumerate(bands):
        m, excl = effmass.kp_sum(E, V, d2[i], n)
        assert excl == [n]
        fd = np.zeros((3, 3))
        for a in range(3):
            for b in range(3):
                ea, eb = np.eye(3)[a] * h, np.eye(3)[b] * h
                fd[a, b] = (En(ea + eb)[n] - En(ea - eb)[n] - En(-ea + eb)[n] + En(-ea - eb)[n]) / (4 * h * h)
        assert np.abs(m - fd).max() < 2e-4 * np.abs(fd).max(), (n, m, fd)


def test_free_electron_mass_is_one():
    wl = _cell()
    cell = copy.copy(wl.cell)
    cell.vloc = np.zeros_like(cell.vloc)
    sp = copy.copy(cell.species_list[0]); sp.lam = np.zeros_like(sp.lam)
    cell.species_list = [sp]
    kp = wl.kpts[0]
    H = refmath.dense_h(cell, kp)
    E, X = np.linalg.eigh(H)
    V = [X.conj().T @ np.diag(kp.K[:, a]) @ X for a in range(3)]
    for n in range(4):
        m, _ = effmass.kp_sum(E, V, np.eye(3), n)
        assert np.abs(m - np.eye(3)).max() < 1e-10
