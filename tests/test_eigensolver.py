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


def test_plane_wave_start():
    """plane_wave_start: when H does not couple the m lowest-kinetic plane
    waves to the others, its start vectors are exact eigenvectors (residual
    at rounding level) and ChFSI stops at once; on the fcc H_GG' the start
    plus ChFSI gives the dense eigenvalues, with fewer H applies than a
    random start; the wanted/uniform buffer degree rules reach the same
    eigenpairs."""
    rng = np.random.default_rng(5)
    n, m, rows = 300, 40, 12
    kin = rng.uniform(0, 30, n)
    low = np.argsort(kin, kind="stable")[:m]
    rest = np.setdiff1d(np.arange(n), low)
    Hd = np.diag(kin).astype(complex)
    for blk in (low, rest):
        A = (rng.standard_normal((len(blk), len(blk))) + 1j * rng.standard_normal((len(blk), len(blk)))) * 0.3
        Hd[np.ix_(blk, blk)] += A + A.conj().T
    H = torch.from_numpy(Hd)
    apply = lambda V, HV: HV.copy_(V @ H.T)          # noqa: E731
    X0 = eigensolver.plane_wave_start(apply, torch.from_numpy(kin), rows, m=m, work_dtype=torch.complex128)
    Hx = X0 @ H.T
    e = torch.real((X0.conj() * Hx).sum(1))
    assert float(torch.linalg.norm(Hx - e[:, None] * X0, dim=1).max()) < 1e-11
    assert np.allclose(np.sort(e.numpy()), np.linalg.eigvalsh(Hd[np.ix_(low, low)])[:rows], atol=1e-11)
    wl = workload("si2_fcc")
    kp = wl.kpts[0]
    Hf = refmath.dense_h(wl.cell, kp)
    Ht = torch.from_numpy(Hf)
    kt = torch.from_numpy(kp.kin)
    ref = np.linalg.eigvalsh(Hf)[:8]
    counts = {}
    for name, kw in (("random", {}), ("pw", {}), ("pw wanted", {"buffer_degree": "wanted"}),
                     ("pw uniform", {"buffer_degree": "wanted", "uniform": True})):
        cnt = [0]

        def app(V, HV):
            cnt[0] += V.shape[0]
            HV.copy_(V @ Ht.T)
        if name == "random":
            X0 = eigensolver.random_start(12, kp.npw, kt, seed=2)
        else:
            X0 = eigensolver.plane_wave_start(app, kt, 12, work_dtype=torch.complex128, eigh=eigensolver.eigh_host)
        res = eigensolver.chfsi(app, X0, kt, nconv=8, tol=1e-9, eigh=eigensolver.eigh_host, **kw)
        assert res.converged == 8, name
        assert np.abs(res.E[:8].numpy() - ref).max() < 1e-11, name
        counts[name] = cnt[0]
    assert counts["pw"] < counts["random"], counts
