"""Effective masses on the device (the 8-atom cubic cell, ecut 50 Ry,
npw 6451, 200 bands at Gamma, bands n = 4, 5):
* FP64 k.p masses equal the oracle's k.p masses on the same eigenpairs;
* the mixed path (FP32 H|psi>, dH/dk; FP64 sums) is within 1e-3 relative;
* free electrons give 1/m* = identity."""
import copy

import numpy as np
import pytest
import torch

import oracle
from conftest import workload
from effmass_amd import effmass, eigensolver, synth
from effmass_amd.hamiltonian import Hamiltonian

pytestmark = pytest.mark.gpu
BANDS = [4, 5]
_CACHE = {}


def solve(prec):
    if prec in _CACHE:
        return _CACHE[prec]
    wl = workload("si8_cubic", nb=2)
    cell, kp = wl.cell, wl.kpts[0]
    H = Hamiltonian(cell)
    H.load_kpoint(0, kp)
    kin = torch.from_numpy(kp.kin).cuda()
    nb = 208
    X0 = eigensolver.random_start(nb, kp.npw, kin, seed=11, device="cuda")
    op = H.at_k(0)
    wdt = torch.complex128 if prec == "fp64" else torch.complex64
    tol = 1e-8 if prec == "fp64" else 2e-5
    res = eigensolver.lobpcg(lambda V, HV: op(V, HV, prec=prec, overwrite=True), X0, kin, nconv=200,
                             tol=tol, maxiter=600, work_dtype=wdt)
    print(f"{prec}: LOBPCG {res.iters} it, converged {res.converged}/200", flush=True)
    X = res.X[:200].contiguous()
    m = effmass.effective_mass(H, 0, kp, cell.A, X, res.E[:200], BANDS, prec=prec)
    _CACHE[prec] = (wl, H, X, res.E[:200].cpu().numpy(), m, res)
    return _CACHE[prec]


def test_fp64_mass_matches_oracle_on_same_eigenpairs():
    wl, H, X, E, m, res = solve("fp64")
    assert res.converged == 200
    cell, kp = wl.cell, wl.kpts[0]
    Xh = X.cpu().numpy()
    D = oracle.dhdk(cell, kp, Xh)
    V = [Xh.conj() @ d.T for d in D]
    d2 = np.zeros((len(BANDS), 3, 3))
    delta = 1e-3
    for b in range(3):
        dk = np.zeros(3); dk[b] = delta
        Dp = oracle.dhdk(cell, effmass.shifted_kpoint(kp, cell.A, dk), Xh[BANDS])
        Dm = oracle.dhdk(cell, effmass.shifted_kpoint(kp, cell.A, -dk), Xh[BANDS])
        for a in range(3):
            d2[:, a, b] = np.real(np.sum(Xh[BANDS].conj() * (Dp[a] - Dm[a]), 1)) / (2 * delta)
    d2 = 0.5 * (d2 + d2.transpose(0, 2, 1))
    for i, n in enumerate(BANDS):
        ref, _ = effmass.kp_sum(E, V, d2[i], n)
        err = np.abs(m.inv_mass[i] - ref).max() / np.abs(ref).max()
        print(f"band {n}: 1/m* diag {np.diag(m.inv_mass[i])}, masses {m.masses()[i]}, vs oracle {err:.2e}",
              flush=True)
        assert err < 1e-9


def test_mixed_mass_within_1e3_of_fp64():
    m64 = solve("fp64")[4]
    m32 = solve("mixed")[4]
    for i, n in enumerate(BANDS):
        rel = np.abs(m32.inv_mass[i] - m64.inv_mass[i]).max() / np.abs(m64.inv_mass[i]).max()
        mrel = np.abs(m32.masses()[i] - m64.masses()[i]) / np.abs(m64.masses()[i])
        print(f"band {n}: mixed vs fp64 1/m* {rel:.2e}, m* {mrel.max():.2e}", flush=True)
        assert mrel.max() < 1e-3


def test_free_electron_mass_on_device():
    wl = synth.custom(*synth.fcc_primitive(10.26), (24, 24, 24), 12.0,
                      [np.array([0.11, 0.23, -0.17])], 6, seed=21)
    cell = copy.copy(wl.cell)
    cell.vloc = np.zeros_like(cell.vloc)
    sp = copy.copy(cell.species_list[0]); sp.lam = np.zeros_like(sp.lam)
    cell.species_list = [sp]
    kp = wl.kpts[0]
    H = Hamiltonian(cell)
    H.load_kpoint(0, kp)
    kin = torch.from_numpy(kp.kin).cuda()
    X0 = eigensolver.random_start(10, kp.npw, kin, seed=3, device="cuda")
    op = H.at_k(0)
    res = eigensolver.lobpcg(lambda V, HV: op(V, HV, overwrite=True), X0, kin, nconv=8, tol=1e-10)
    m = effmass.effective_mass(H, 0, kp, cell.A, res.X[:8].contiguous(), res.E[:8], [0, 1, 2])
    assert np.abs(m.inv_mass - np.eye(3)[None]).max() < 1e-8
    H.close()


def test_device_kp_sum_equals_host_sum():
    """effm_kp_mass (FP64 sum and denominators on the device) equals the
    host kp_sum on the same random velocity matrices, eigenvalues with a
    degenerate pair and d2h, to 1e-13; both orders of the symmetrisation."""
    rng = np.random.default_rng(3)
    nb = 77
    E = np.sort(rng.uniform(-1, 1, nb))
    E[10] = E[11] = 0.1234                                     # a degenerate pair
    V = []
    for _ in range(3):
        M = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
        V.append(M + M.conj().T)                               # Hermitian velocity matrices
    bands = [0, 10, 11, 40, 76]
    d2 = rng.standard_normal((len(bands), 3, 3))
    wl = synth.custom(*synth.fcc_primitive(10.26), (24, 24, 24), 12.0, [np.zeros(3)], 2, seed=1)
    H = Hamiltonian(wl.cell)
    got = H.kp_mass(torch.from_numpy(E).cuda(), [torch.from_numpy(v).cuda() for v in V], bands,
                    torch.from_numpy(d2).cuda(), 1e-6).cpu().numpy()
    H.close()
    for i, n in enumerate(bands):
        ref, excl = effmass.kp_sum(E, V, d2[i], n, 1e-6)
        assert np.abs(got[i] - ref).max() <= 1e-13 * np.abs(ref).max(), n
        if n in (10, 11):
            assert excl == [10, 11]


def test_velocity_columns_equal_full_matrices():
    """effective_mass reads only the rows and columns of the asked bands:
    velocity_columns (dH/dk on those states only, rows as conjugate columns)
    equals the full velocity matrices there to 1e-12 of their largest entry
    (FP64, 200 states, bands 4 and 5)."""
    wl, H, X, E, m, res = solve("fp64")
    full = effmass.velocity_matrices(H, 0, X, "fp64")
    cols = effmass.velocity_columns(H, 0, X, BANDS, "fp64")
    for a in range(3):
        F, Cm = full[a].cpu().numpy(), cols[a].cpu().numpy()
        scale = np.abs(F).max()
        for n in BANDS:
            ec = np.abs(Cm[:, n] - F[:, n]).max() / scale
            er = np.abs(Cm[n, :] - F[n, :]).max() / scale
            print(f"dir {a} band {n}: column {ec:.1e}, row {er:.1e}", flush=True)
            assert ec <= 1e-12 and er <= 1e-12
