"""The engine on real PWmat / SG15 inputs (fixtures copied from the
reference's test data by tests/golden/make_pwmat_fixtures.py):

* FP64 device H|psi> equals the oracle on the 8-atom Si and 4-atom AlN cells
  (real UPF projectors, two species, a 20x20x32 grid);
* the lowest eigenvalues of the device H (LOBPCG, FP64 Rayleigh-Ritz)
  match PWmat's OUT.EIGEN within the reference's own tolerance of
  2e-5 Ha (test/H_psi/test_hpsi_eigen.cpp:37)."""
import os

import numpy as np
import pytest
import torch

import oracle
from effmass_amd import eigensolver, pwmat
from effmass_amd.hamiltonian import Hamiltonian

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pwmat")
CASES = {
    "si8": ("si8_atom.config", "si8_IN.VR", {14: "Si.SG15.PBE.UPF"}, "si8_OUT2_k0.GKK", "si8_OUT2.EIGEN"),
    "aln": ("aln_atom.config", "aln_OUT.VR", {7: "N.SG15.PBE.UPF", 13: "Al.SG15.PBE.UPF"},
            "aln_OUT_k0.GKK", "aln_OUT.EIGEN"),
}


def load(case):
    ac, vr, upfs, gkk, eig = CASES[case]
    g = pwmat.GKKFile(os.path.join(G, gkk))
    cell = pwmat.load_cell(os.path.join(G, ac), os.path.join(G, vr),
                           {z: os.path.join(G, f) for z, f in upfs.items()}, n=g.meta.n)
    return cell, g.kpoint(0), pwmat.read_eigen(os.path.join(G, eig))["eig"][0, 0]


@pytest.mark.parametrize("case", ["si8", "aln"])
def test_fixture_cell_parity(case):
    cell, kp, _ = load(case)
    H = Hamiltonian(cell)
    H.load_kpoint(0, kp)
    psi = eigensolver.random_start(6, kp.npw, torch.from_numpy(kp.kin), seed=4).numpy()
    d = torch.from_numpy(psi).cuda()
    h = torch.zeros_like(d)
    H.at_k(0)(d, h)
    torch.cuda.synchronize()
    ref = oracle.hpsi(cell, kp, psi)
    got = h.cpu().numpy()
    e = (np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)).max()
    print(f"{case} fixture parity {e:.2e}", flush=True)
    assert e <= 1e-12
    H.close()


@pytest.mark.parametrize("case,nev", [("si8", 26), ("aln", 20)])
def test_eigenvalues_vs_pwmat(case, nev):
    cell, kp, ref = load(case)
    H = Hamiltonian(cell)
    H.load_kpoint(0, kp)
    kin = torch.from_numpy(kp.kin).cuda()
    nb = nev + 8
    X0 = eigensolver.random_start(nb, kp.npw, kin, seed=7, device="cuda")
    op = H.at_k(0)
    res = eigensolver.lobpcg(lambda V, HV: op(V, HV, overwrite=True), X0, kin, nconv=nev, tol=1e-6,
                             maxiter=400)
    E = res.E[:nev].cpu().numpy()
    dE = np.abs(E - ref[:nev])
    print(f"{case}: {res.iters} iterations, max |E - E_PWmat| = {dE.max():.2e} Ha", flush=True)
    assert res.converged == nev
    assert dE.max() < 2e-5
    H.close()
