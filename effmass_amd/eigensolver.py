"""Block eigensolver for the lowest bands of H at one k-point.

The reference has no eigensolver (docs/roadmap.md:5-7); its tests take
eigenvectors from PWmat's OUT.WG.  This is the caller of the batched
H|psi> that the effective-mass and serving layers need: LOBPCG with
FP64 Cholesky orthonormalisation and Rayleigh-Ritz on the device, and the
Teter-Payne-Allan preconditioner in the kinetic energy.  The H|psi>
application may run in FP64 or in the mixed path (FP32 FFT / GEMM); the
subspace algebra stays FP64 either way.

Vectors are band-major rows: X[b, g].  Overlaps are conj(X) @ Y^T.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass
class EigResult:
    E: torch.Tensor          # [nb] float64, ascending
    X: torch.Tensor          # [nb, npw] complex128, orthonormal rows
    HX: torch.Tensor         # H X (same precision as X)
    resid: torch.Tensor      # [nb] residual norms ||HX - E X||
    iters: int
    converged: int


def _gram(A, B):
    return A.conj() @ B.T


def _chol_orth(V, HV=None, tries=3):
    """Orthonormalise the rows of V by Cholesky of the FP64 Gram matrix; HV
    (if given) is transformed alike.  A growing diagonal shift handles a
    near-singular basis."""
    G = _gram(V, V)
    G = 0.5 * (G + G.conj().T)
    eye = torch.eye(G.shape[0], dtype=G.dtype, device=G.device)
    shift = 0.0
    for _ in range(tries):
        L, info = torch.linalg.cholesky_ex(G + shift * eye)
        if int(info) == 0:
            break
        shift = (shift * 100) if shift else 1e-14 * float(G.diagonal().real.max())
    else:
        raise RuntimeError("basis is numerically rank deficient")
    # rows: V_new = L^{-1} conj? With G = conj(V) V^T = L L^H the rows of
    # L^{-1}.conj() V are orthonormal: Gram = conj(L^-1) G L^-T = I.
    Li = torch.linalg.solve_triangular(L, eye, upper=False).conj()
    Vn = Li @ V
    return (Vn, Li @ HV) if HV is not None else (Vn, None)


def tpa_precond(kin, E):
    """Teter-Payne-Allan: P(x) = (27+18x+12x^2+8x^3)/(27+18x+12x^2+8x^3+16x^4),
    x = T(G)/T_band, T_band ~ max(E_b, 1) per band."""
    t = torch.clamp(E.abs(), min=1.0)[:, None]
    x = kin[None, :] / t
    p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
    return p / (p + 16.0 * x ** 4)


def lobpcg(apply, X0: torch.Tensor, kin: torch.Tensor, nconv: int | None = None, tol: float = 1e-8,
           maxiter: int = 300, work_dtype=None) -> EigResult:
    """Lowest nb = X0.shape[0] eigenpairs.  apply(X, HX) writes H X into HX
    (X, HX in work_dtype, default X0's dtype).  Converged when the first
    nconv (default all) residual norms are below tol."""
    dev = X0.device
    cdt = torch.complex128
    wdt = work_dtype or X0.dtype
    nb = X0.shape[0]
    nconv = nb if nconv is None else nconv
    kin = kin.to(device=dev, dtype=torch.float64)

    def H(V):
        Vw = V.to(wdt).contiguous()
        HV = torch.empty_like(Vw)
        apply(Vw, HV)
        return HV.to(cdt)

    X, _ = _chol_orth(X0.to(cdt))
    HX = H(X)
    Hs = _gram(X, HX)
    e, C = torch.linalg.eigh(0.5 * (Hs + Hs.conj().T))
    X = C.T @ X
    HX = C.T @ HX
    P = HP = None
    it = 0
    conv = 0
    for it in range(1, maxiter + 1):
        R = HX - e[:, None] * X
        rn = torch.linalg.norm(R, dim=1)
        conv = int((rn[:nconv] < tol).sum())
        if conv == nconv:
            break
        W = R * tpa_precond(kin, e)
        # block of the unconverged directions only; W orthogonal to X
        W = W - (_gram(X, W)).T @ X
        blocks = [X, W] + ([P] if P is not None else [])
        S = torch.cat(blocks, 0)
        S, _ = _chol_orth(S)
        HS = H(S)
        Hs = _gram(S, HS)
        ev, Cv = torch.linalg.eigh(0.5 * (Hs + Hs.conj().T))
        Cn = Cv[:, :nb]
        Xn = Cn.T @ S
        HXn = Cn.T @ HS
        # conjugate direction: the part of the new X outside the old X
        Cp = Cn.clone()
        Cp[:nb] = 0
        P = Cp.T @ S
        HP = Cp.T @ HS
        P, HP = _chol_orth(P, HP) if torch.linalg.norm(P) > 0 else (None, None)
        X, HX, e = Xn, HXn, ev[:nb]
    R = HX - e[:, None] * X
    rn = torch.linalg.norm(R, dim=1)
    return EigResult(E=e, X=X, HX=HX, resid=rn, iters=it, converged=int((rn[:nconv] < tol).sum()))


def random_start(nb, npw, kin, seed=0, device="cpu"):
    """Deterministic start vectors: complex normal weighted by 1/(1+T(G))."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    re = torch.randn(nb, npw, generator=g, dtype=torch.float64)
    im = torch.randn(nb, npw, generator=g, dtype=torch.float64)
    w = 1.0 / (1.0 + kin.detach().cpu().to(torch.float64))
    return ((re + 1j * im) * w[None, :]).to(device)
