"""d2H/dk_a dk_b |psi> (the Hessian operator the reference declares,
src/H_psi/hamiltonian.cppm:267-306, with an empty body,
src/H_psi/hamiltonian_hessian.cpp:48-50), as central differences of the
device dH/dk over k:
* free electrons: exactly delta_ab psi (the kinetic dH/dk_a = K_a is linear
  in k, so the difference has no truncation error);
* every component equals the same central difference of the oracle's
  dH/dk (FP64, 1e-9), and the mixed path within 1e-3 of the diagonal's norm;
* the diagonal of <psi|d2H|psi> equals the expectations the k.p mass uses
  (effmass.d2h_expect), the off-diagonal ones within the O(delta^2) of the
  two difference directions;
* accumulate semantics, skipped components and host arrays."""
import copy

import numpy as np
import pytest
import torch

import oracle
from conftest import workload
from effmass_amd import effmass, synth
from effmass_amd.hamiltonian import Hamiltonian, HessHPsi, shifted_kpoint

pytestmark = pytest.mark.gpu
COMP = {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}
DELTA = 1e-3


def _oracle_hess(cell, kp, psi, name):
    a, b = COMP[name]
    dk = np.zeros(3)
    dk[b] = DELTA
    dp = oracle.dhdk(cell, shifted_kpoint(kp, cell.A, dk), psi)[a]
    dm = oracle.dhdk(cell, shifted_kpoint(kp, cell.A, -dk), psi)[a]
    return (dp - dm) / (2 * DELTA)


def _rel(got, ref):
    return float((np.linalg.norm(got - ref, axis=-1) / np.linalg.norm(ref, axis=-1)).max())


def test_free_electrons_give_identity():
    wl = synth.custom(*synth.fcc_primitive(10.26), (24, 24, 24), 12.0, [np.array([0.11, 0.23, -0.17])], 3,
                      seed=5)
    cell = copy.copy(wl.cell)
    cell.vloc = np.zeros_like(cell.vloc)
    sp = copy.copy(cell.species_list[0]); sp.lam = np.zeros_like(sp.lam)
    cell.species_list = [sp]
    H = Hamiltonian(cell)
    H.load_kpoint(0, wl.kpts[0])
    psi = torch.from_numpy(wl.psi[0]).cuda()
    out = HessHPsi(**{n: torch.zeros_like(psi) for n in COMP})
    H.hessian(DELTA).at_k(0)(psi, out)
    for n, (a, b) in COMP.items():
        ref = psi if a == b else torch.zeros_like(psi)
        err = float((getattr(out, n) - ref).abs().max())
        print(f"free electrons {n}: max |err| {err:.2e}", flush=True)
        assert err < 1e-10
    H.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_hessian_equals_oracle_difference(prec):
    wl = workload("si2_fcc")
    cell, kp = wl.cell, wl.kpts[0]
    H = Hamiltonian(cell)
    H.load_kpoint(0, kp)
    cdt = torch.complex128 if prec == "fp64" else torch.complex64
    psi = torch.from_numpy(wl.psi[0]).to(cdt).cuda()
    out = HessHPsi(**{n: torch.zeros_like(psi) for n in COMP})
    H.hessian(DELTA).at_k(0)(psi, out, prec=prec)
    # FP64: each component relative to its own norm; mixed: relative to the
    # norm of the diagonal component, since an FP32 difference over 2 delta
    # carries ~1e-7 |dH/dk psi| / 2 delta whatever the component's size
    refs = {n: _oracle_hess(cell, kp, wl.psi[0], n) for n in COMP}
    diag = np.linalg.norm(refs["xx"], axis=-1)
    for n in COMP:
        got = getattr(out, n).cpu().numpy().astype(np.complex128)
        if prec == "fp64":
            e = _rel(got, refs[n])
            assert e <= 1e-9, (n, e)
        else:
            e = float((np.linalg.norm(got - refs[n], axis=-1) / diag).max())
            assert e <= 1e-3, (n, e)
        print(f"{prec} {n}: err vs oracle difference {e:.2e}", flush=True)
    assert sorted(H._npw) == [0]                      # the scratch slot was released
    H.close()


def test_expectations_match_kp_d2h():
    wl = workload("si2_fcc")
    cell, kp = wl.cell, wl.kpts[0]
    H = Hamiltonian(cell)
    H.load_kpoint(0, kp)
    psi = torch.from_numpy(wl.psi[0]).cuda()
    out = HessHPsi(**{n: torch.zeros_like(psi) for n in COMP})
    H.hessian(DELTA).at_k(0)(psi, out)
    bands = list(range(psi.shape[0]))
    d2 = effmass.d2h_expect(H, kp, cell.A, psi, bands, delta=DELTA).cpu().numpy()
    for n, (a, b) in COMP.items():
        e = (psi.conj() * getattr(out, n)).sum(1).real.cpu().numpy()
        diff = np.abs(e - d2[:, a, b]).max()
        scale = np.abs(d2).max()
        print(f"<psi|d2H_{n}|psi> vs d2h_expect: {diff:.2e} (scale {scale:.2e})", flush=True)
        assert diff <= (1e-10 if a == b else 1e-5) * scale
    H.close()


def test_accumulate_skip_and_host_arrays():
    wl = workload("si2_fcc")
    cell, kp = wl.cell, wl.kpts[0]
    H = Hamiltonian(cell)
    H.load_kpoint(0, kp)
    psi = wl.psi[0][:3].copy()
    hc = H.hessian(DELTA).at_k(0)
    ref = HessHPsi(xx=torch.zeros(psi.shape, dtype=torch.complex128, device="cuda"),
                   yz=torch.zeros(psi.shape, dtype=torch.complex128, device="cuda"))
    hc(torch.from_numpy(psi).cuda(), ref)
    xx = np.full(psi.shape, 0.5 - 0.25j)
    yz = np.full(psi.shape, -1.0 + 2.0j)
    hc(psi, HessHPsi(xx=xx, yz=yz))                   # host arrays, xy/xz/yy/zz skipped, accumulate
    assert np.abs(xx - (0.5 - 0.25j) - ref.xx.cpu().numpy()).max() < 1e-12
    assert np.abs(yz - (-1.0 + 2.0j) - ref.yz.cpu().numpy()).max() < 1e-12
    hc(psi, HessHPsi(xx=xx), overwrite=True)
    assert np.abs(xx - ref.xx.cpu().numpy()).max() < 1e-12
    H.close()
