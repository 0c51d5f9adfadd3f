"""LOBPCG (effmass_amd/eigensolver.py) against dense diagonalisation: a
random Hermitian matrix with a kinetic-like diagonal (eigenvalues fixed by
numpy.linalg.eigh), and the 2-atom fcc H_GG' built from the mathematics."""
import numpy as np
import torch

import refmath
from conftest import workload
from effmass_amd import eigensolver


def test_lobpcg_random_hermitian():
    rng = np.random.default_rng(3)
    n = 400
    d = np.sort(rng.uniform(0, 20, n))
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 0.05
    Hd = np.diag(d) + A + A.conj().T
    H = torch.from_numpy(Hd)
    kin = torch.from_numpy(d)
    X0 = eigensolver.random_start(12, n, kin, seed=1)
    res = eigensolver.lobpcg(lambda V, HV: HV.copy_(V @ H.T), X0, kin, nconv=10, tol=1e-9)
    ref = np.linalg.eigvalsh(Hd)[:10]
    assert res.converged == 10
    assert np.abs(res.E[:10].numpy() - ref).max() < 1e-12 * max(1, abs(ref).max())
    # orthonormal rows, Ritz pairs
    G = (res.X.conj() @ res.X.T).numpy()
    assert np.abs(G - np.eye(12)).max() < 1e-12


def test_lobpcg_dense_fcc_h():
    wl = workload("si2_fcc")
    kp = wl.kpts[0]
    Hd = refmath.dense_h(wl.cell, kp)
    H = torch.from_numpy(Hd)
    kin = torch.from_numpy(kp.kin)
    X0 = eigensolver.random_start(10, kp.npw, kin, seed=2)
    res = eigensolver.lobpcg(lambda V, HV: HV.copy_(V @ H.T), X0, kin, nconv=8, tol=1e-9)
    ref = np.linalg.eigvalsh(Hd)[:8]
    assert res.converged == 8
    assert np.abs(res.E[:8].numpy() - ref).max() < 1e-11


def test_chfsi_random_hermitian_and_fcc():
    """Chebyshev-filtered subspace iteration: same eigenpairs as dense
    diagonalisation, on the random matrix and the 2-atom fcc H_GG'."""
    rng = np.random.default_rng(3)
    n = 400
    d = np.sort(rng.uniform(0, 20, n))
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 0.05
    Hd = np.diag(d) + A + A.conj().T
    H = torch.from_numpy(Hd)
    kin = torch.from_numpy(d)
    X0 = eigensolver.random_start(14, n, kin, seed=1)
    res = eigensolver.chfsi(lambda V, HV: HV.copy_(V @ H.T), X0, kin, nconv=10, tol=1e-9)
    ref = np.linalg.eigvalsh(Hd)[:10]
    assert res.converged == 10
    assert np.abs(res.E[:10].numpy() - ref).max() < 1e-12 * max(1, abs(ref).max())
    G = (res.X.conj() @ res.X.T).numpy()
    assert np.abs(G - np.eye(14)).max() < 1e-12
    wl = workload("si2_fcc")
    kp = wl.kpts[0]
    Hf = refmath.dense_h(wl.cell, kp)
    Ht = torch.from_numpy(Hf)
    X0 = eigensolver.random_start(12, kp.npw, torch.from_numpy(kp.kin), seed=2)
    res = eigensolver.chfsi(lambda V, HV: HV.copy_(V @ Ht.T), X0, torch.from_numpy(kp.kin), nconv=8, tol=1e-9,
                            eigh=eigensolver.eigh_host)
    assert res.converged == 8
    assert np.abs(res.E[:8].numpy() - np.linalg.eigvalsh(Hf)[:8]).max() < 1e-11
