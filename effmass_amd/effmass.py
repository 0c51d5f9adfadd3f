"""k.p effective-mass tensor on top of the device H|psi> and dH/dk|psi>.

For a band n at k (Hartree atomic units, hbar = m_e = 1):

  (1/m*)_ab = d^2 E_n / dk_a dk_b
            = <n| d^2H/dk_a dk_b |n>
              + 2 Re sum_{m != n} <n|dH/dk_a|m><m|dH/dk_b|n> / (E_n - E_m)

d^2H/dk_a dk_b is delta_ab (kinetic) plus the second k-derivative of the
nonlocal term.  The reference declares that Hessian but leaves it empty
(src/H_psi/hamiltonian_hessian.cpp:48-50); here <n|d^2H|n> is the central
difference in k of the analytic <n|dH/dk_a(k +- delta e_b)|n>
(src/H_psi/hamiltonian_gradient.cpp:116-294 on the device), which is
exact for the kinetic part and O(delta^2) for the nonlocal part.

Velocity matrices, the sum over states and the (E_n - E_m) denominators are
FP64 in both precisions and stay on the device (effm_overlap for the
matrices, effm_kp_mass for the sum); only the H|psi> / dH/dk applications
follow `prec`.  kp_sum below is the same sum on the host, for dense-matrix
checks on small cells.
States m within `degen_tol` of E_n are left out of the sum (their
contribution is undefined for a degenerate level; the result is then the
mass of the level projected out of its degenerate partners).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import synth
from .hamiltonian import GradHPsi, Hamiltonian
from .hamiltonian import shifted_kpoint as _shifted_kpoint


@dataclass
class MassResult:
    bands: list
    inv_mass: np.ndarray        # [len(bands), 3, 3]  (1/m*)_ab
    d2h: np.ndarray             # [len(bands), 3, 3]  <n|d^2H/dk_a dk_b|n>
    E: np.ndarray               # [nb] eigenvalues used
    excluded: list              # per band: indices left out as degenerate

    def masses(self):
        """Principal effective masses (inverse eigenvalues of 1/m*)."""
        return np.array([1.0 / np.linalg.eigvalsh(0.5 * (m + m.T)) for m in self.inv_mass])


def shifted_kpoint(kp: synth.KPoint, A: np.ndarray, dk_cart) -> synth.KPoint:
    """Same G set at k + dk (Cartesian): K + dk, |K|^2/2, k_frac + dk.A^T/2pi."""
    return _shifted_kpoint(kp, A, dk_cart)


def kp_sum(E, V, d2h, n, degen_tol=1e-6):
    """(1/m*)_ab for band n from eigenvalues E [nb], velocity matrices
    V[a][m, l] = <m|dH/dk_a|l> (3 x nb x nb) and <n|d^2H|n> (3x3), on the
    host (numpy): the dense-matrix form of effm_kp_mass."""
    E = np.asarray(E, dtype=np.float64)
    d = E[n] - E
    keep = np.abs(d) > degen_tol
    keep[n] = False
    M = np.array(d2h, dtype=np.float64).copy()
    for a in range(3):
        for b in range(3):
            t = V[a][n, keep] * V[b][keep, n] / d[keep]
            M[a, b] += 2.0 * np.real(t.sum())
    return 0.5 * (M + M.T), np.nonzero(~keep)[0].tolist()


def velocity_matrices(H: Hamiltonian, ikpt: int, X: torch.Tensor, prec="fp64"):
    """V[a] = <m|dH/dk_a|n> (FP64 complex128 device tensors, nb x nb) for
    the rows of X."""
    Xp = X.to(torch.complex128 if prec == "fp64" else torch.complex64).contiguous()
    D = [torch.zeros_like(Xp) for _ in range(3)]
    H.gradient().at_k(ikpt)(Xp, GradHPsi(*D), prec=prec, overwrite=True)
    X64 = X.to(torch.complex128).contiguous()
    # FP64 split-K products on the engine (a long-K complex128 matmul on
    # the BLAS library runs at a fraction of the MFMA rate)
    return [H.overlap(X64, d.to(torch.complex128)) for d in D]


def velocity_columns(H: Hamiltonian, ikpt: int, X: torch.Tensor, bands, prec="fp64"):
    """The entries of the velocity matrices the k.p sum of `bands` reads:
    V[a][m][n] = <m|dH/dk_a|n> for n in bands (dH/dk applied to those
    states only, FP64 overlaps with all rows of X) and the rows V[a][n][m] =
    conj(V[a][m][n]) (dH/dk_a is Hermitian); complex128 [nb, nb] device
    tensors, zero elsewhere.  For a few bands of many states this applies
    dH/dk to len(bands) states instead of nb."""
    nb = X.shape[0]
    idx = torch.as_tensor([int(n) for n in bands], dtype=torch.long, device=X.device)
    Xs = X[idx].to(torch.complex128 if prec == "fp64" else torch.complex64).contiguous()
    D = [torch.zeros_like(Xs) for _ in range(3)]
    H.gradient().at_k(ikpt)(Xs, GradHPsi(*D), prec=prec, overwrite=True)
    X64 = X.to(torch.complex128).contiguous()
    V = []
    for d in D:
        cols = H.overlap(X64, d.to(torch.complex128))          # [nb, len(bands)]: <m|dH/dk_a|n>
        Va = torch.zeros((nb, nb), dtype=torch.complex128, device=X.device)
        Va[:, idx] = cols
        Va[idx, :] = cols.conj().T
        V.append(Va)
    return V


def d2h_expect(H: Hamiltonian, kp: synth.KPoint, A, X: torch.Tensor, bands, prec="fp64",
               delta=1e-3, slot=None):
    """<n| d^2H / dk_a dk_b |n> for the bands (rows of X) by central
    differences of the analytic gradient; uses one scratch k-point slot
    (default: one past the highest loaded)."""
    if slot is None:
        slot = max(H._npw.keys(), default=-1) + 1
    Xb = X[bands].to(torch.complex128 if prec == "fp64" else torch.complex64).contiguous()
    X64 = X[bands].to(torch.complex128)
    out = torch.zeros((len(bands), 3, 3), dtype=torch.float64, device=X.device)
    for b in range(3):
        vals = []
        for sgn in (+1.0, -1.0):
            dk = np.zeros(3)
            dk[b] = sgn * delta
            H.load_kpoint(slot, shifted_kpoint(kp, A, dk))
            D = [torch.zeros_like(Xb) for _ in range(3)]
            H.gradient().at_k(slot)(Xb, GradHPsi(*D), prec=prec, overwrite=True)
            vals.append(torch.stack([(X64.conj() * d.to(torch.complex128)).sum(1).real for d in D],
                                    dim=1))                  # [nbands, a], FP64 on the device
        out[:, :, b] = (vals[0] - vals[1]) / (2.0 * delta)
    H.release_kpoint(slot)
    return 0.5 * (out + out.transpose(1, 2))


def effective_mass(H: Hamiltonian, ikpt: int, kp: synth.KPoint, A, X: torch.Tensor, E, bands,
                   prec="fp64", delta=1e-3, degen_tol=1e-6) -> MassResult:
    """(1/m*) for `bands` from the eigenpairs (X rows, E) of H at ikpt."""
    Ed = (E if torch.is_tensor(E) else torch.as_tensor(np.asarray(E, dtype=np.float64))).to(
        device=X.device, dtype=torch.float64)
    # only the rows / columns of the asked bands enter the sum
    V = velocity_columns(H, ikpt, X, bands, prec)
    d2 = d2h_expect(H, kp, A, X, bands, prec, delta)
    inv = H.kp_mass(Ed, V, bands, d2, degen_tol).cpu().numpy()
    En = Ed.cpu().numpy()
    excl = []
    for n in bands:
        keep = np.abs(En[n] - En) > degen_tol
        keep[n] = False
        excl.append(np.nonzero(~keep)[0].tolist())
    return MassResult(bands=list(bands), inv_mass=inv, d2h=d2.cpu().numpy(), E=En, excluded=excl)
