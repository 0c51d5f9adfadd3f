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


def _chol_orth(V, HV=None, tries=3, gram=None):
    """Orthonormalise the rows of V by Cholesky of the FP64 Gram matrix; HV
    (if given) is transformed alike.  A growing diagonal shift handles a
    near-singular basis."""
    G = (gram or _gram)(V, V)
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


def _rr(A, B):
    """Generalized Rayleigh-Ritz A c = lambda B c (FP64, Hermitian, B > 0)
    through the Cholesky factor of B; returns (lambda ascending, C) with the
    eigenvectors in the columns of C, or None if B is not positive definite."""
    A = 0.5 * (A + A.conj().T)
    B = 0.5 * (B + B.conj().T)
    L, info = torch.linalg.cholesky_ex(B)
    if int(info) != 0:
        return None
    eye = torch.eye(B.shape[0], dtype=B.dtype, device=B.device)
    Li = torch.linalg.solve_triangular(L, eye, upper=False)
    Ap = Li @ A @ Li.conj().T
    lam, V = torch.linalg.eigh(0.5 * (Ap + Ap.conj().T))
    return lam, Li.conj().T @ V


def lobpcg(apply, X0: torch.Tensor, kin: torch.Tensor, nconv: int | None = None, tol: float = 1e-8,
           maxiter: int = 300, work_dtype=None, gram=None) -> EigResult:
    """Lowest nb = X0.shape[0] eigenpairs.  apply(X, HX) writes H X into HX
    (X, HX in work_dtype, default X0's dtype).  Converged when the first
    nconv (default all) residual norms are below tol.  gram(A, B) computes
    the FP64 overlaps conj(A) @ B.T (default torch; the engine's split-K
    kernel, Hamiltonian.overlap, is much faster for long rows).

    Each iteration applies H only to the new directions W (residuals of the
    bands not yet converged, preconditioned and orthonormalised against X);
    H X and H P are carried along.  The Rayleigh-Ritz step works on the
    block Gram matrices of [X, W, P] in FP64 (S^H H S and S^H S), solved as
    a generalized problem through a Cholesky factor; if that factor fails
    the conjugate block P is dropped for the iteration."""
    dev = X0.device
    cdt = torch.complex128
    wdt = work_dtype or X0.dtype
    nb = X0.shape[0]
    nconv = nb if nconv is None else nconv
    kin = kin.to(device=dev, dtype=torch.float64)

    gr = gram or _gram

    def H(V):
        Vw = V.to(wdt).contiguous()
        HV = torch.empty_like(Vw)
        apply(Vw, HV)
        return HV.to(cdt)

    X, _ = _chol_orth(X0.to(cdt), gram=gram)
    HX = H(X)
    G0 = gr(X, HX)
    e, C = torch.linalg.eigh(0.5 * (G0 + G0.conj().T))
    X = C.T @ X
    HX = C.T @ HX
    P = HP = None
    it = 0
    for it in range(1, maxiter + 1):
        R = HX - e[:, None] * X
        rn = torch.linalg.norm(R, dim=1)
        conv = int((rn[:nconv] < tol).sum())
        if conv == nconv:
            break
        act = torch.nonzero(rn >= tol).flatten()          # soft locking
        W = R[act] * tpa_precond(kin, e[act])
        W = W - gr(X, W).T @ X
        try:
            W, _ = _chol_orth(W, gram=gram)
        except RuntimeError:
            break                                          # no new directions left
        HW = H(W)
        blocks = [(X, HX), (W, HW)] + ([(P, HP)] if P is not None else [])
        for attempt in range(2):
            n_blk = len(blocks)
            A = [[None] * n_blk for _ in range(n_blk)]
            B = [[None] * n_blk for _ in range(n_blk)]
            for i in range(n_blk):
                for j in range(i, n_blk):
                    Si, HSi = blocks[i]
                    Sj, HSj = blocks[j]
                    # X^H X and X^H H X are computed too (not taken as I and
                    # diag(e)), so rounding drift of X never accumulates
                    A[i][j] = gr(Si, HSj)
                    B[i][j] = gr(Si, Sj)
                    if i != j:
                        A[j][i] = A[i][j].conj().T
                        B[j][i] = B[i][j].conj().T
            sol = _rr(torch.cat([torch.cat(r, 1) for r in A], 0), torch.cat([torch.cat(r, 1) for r in B], 0))
            if sol is not None or len(blocks) == 2:
                break
            blocks = blocks[:2]                            # drop P and retry
        if sol is None:
            break
        lam, Cv = sol
        Cn = Cv[:, :nb]
        sizes = [b[0].shape[0] for b in blocks]
        offs = [0] + list(torch.tensor(sizes).cumsum(0).tolist())
        Xn = sum(Cn[offs[i]:offs[i + 1]].T @ blocks[i][0] for i in range(len(blocks)))
        HXn = sum(Cn[offs[i]:offs[i + 1]].T @ blocks[i][1] for i in range(len(blocks)))
        # conjugate direction: the part of the new X outside the old X
        P = sum(Cn[offs[i]:offs[i + 1]].T @ blocks[i][0] for i in range(1, len(blocks)))
        HP = sum(Cn[offs[i]:offs[i + 1]].T @ blocks[i][1] for i in range(1, len(blocks)))
        try:
            P, HP = _chol_orth(P, HP, gram=gram)
        except RuntimeError:
            P = HP = None
        X, HX, e = Xn, HXn, lam[:nb]
    R = HX - e[:, None] * X
    rn = torch.linalg.norm(R, dim=1)
    return EigResult(E=e, X=X, HX=HX, resid=rn, iters=it, converged=int((rn[:nconv] < tol).sum()))


def _spectral_upper_bound(H, n, dev, steps=12, seed=7):
    """Upper bound of the spectrum of H (Hermitian): a few Lanczos steps on
    one random vector; largest Ritz value + the last off-diagonal."""
    g = torch.Generator(device=dev).manual_seed(seed)
    v = torch.complex(torch.randn(1, n, generator=g, dtype=torch.float64, device=dev),
                      torch.randn(1, n, generator=g, dtype=torch.float64, device=dev))
    v = v / torch.linalg.norm(v)
    v_prev = torch.zeros_like(v)
    beta = 0.0
    alphas, betas = [], []
    for _ in range(steps):
        w = H(v)
        alpha = float(torch.real((v.conj() * w).sum()))
        w = w - alpha * v - beta * v_prev
        beta_new = float(torch.linalg.norm(w))
        alphas.append(alpha)
        betas.append(beta_new)
        if beta_new < 1e-12:
            break
        v_prev, v, beta = v, w / beta_new, beta_new
    T = torch.diag(torch.tensor(alphas, dtype=torch.float64))
    for i in range(len(alphas) - 1):
        T[i, i + 1] = T[i + 1, i] = betas[i]
    return float(torch.linalg.eigvalsh(T).max()) + betas[-1]


def eigh_host(A: torch.Tensor):
    """Hermitian eigendecomposition of a small matrix on the host (LAPACK
    zheevd through numpy), results back on A's device.  Measured against
    torch.linalg.eigh on the device (rocSOLVER, FP64) for n = 40-320:
    0.7-12.5 ms here (copies included) against 1.1-9.6 ms there; the
    serving loop uses this one (its queries' latency below capacity is
    lower with it, at the same saturation throughput)."""
    import numpy as np
    a = A.detach().cpu().numpy()
    w, v = np.linalg.eigh(0.5 * (a + a.conj().T))
    return (torch.from_numpy(w).to(A.device), torch.from_numpy(np.ascontiguousarray(v)).to(A.device))


def chfsi(apply, X0: torch.Tensor, kin: torch.Tensor, nconv: int | None = None, tol: float = 1e-8,
          maxiter: int = 60, work_dtype=None, gram=None, degree: int = 12, eigh=None,
          lincomb=None, adaptive: bool = True, min_degree: int = 4,
          buffer_degree: str = "cap", min_rows: int = 0, uniform: bool = False) -> EigResult:
    """Chebyshev-filtered subspace iteration (Zhou, Saad, Tiago and
    Chelikowsky, J. Comput. Phys. 219 (2006) 172): each outer step filters
    the block with a degree-`degree` Chebyshev polynomial that damps
    [E_top, E_max] (scaled three-term recurrence, `degree` H applies),
    orthonormalises it (Cholesky of the FP64 overlap) and does one
    Rayleigh-Ritz of size nb.  Compared with LOBPCG it trades H applies,
    which the engine makes cheap, for far fewer and smaller dense
    eigenproblems.  Same interface and result as lobpcg(); rows of X0 beyond
    nconv act as a buffer that speeds up convergence of the wanted ones.
    The filter recurrence runs in the work precision (complex64 in the
    mixed path: it only has to damp the unwanted part of the spectrum); the
    orthonormalisation overlaps, the Rayleigh-Ritz matrices and the Ritz
    vectors are FP64.  eigh(A) solves the small Rayleigh-Ritz problems
    (default torch.linalg.eigh; eigh_host runs them on the host).
    lincomb(z, a, y, b, x, c) does z <- a z + b y + c x in place (default:
    three torch ops; Hamiltonian.lincomb does it in one device pass).

    adaptive: each wanted vector (row < nconv) gets its own degree, the one
    that brings its residual r_i to tol if the filter damps the rest of the
    spectrum by |C_m(t_i)| = cosh(m acosh|t_i|), t_i its Ritz value mapped
    onto [-1, 1] by the damped interval (Winkelmann, Springer, Di Napoli,
    ACM TOMS 45 (2019) 21, ChASE), plus one, within [min_degree, degree];
    converged wanted vectors are not filtered; the buffer rows keep
    `degree`.  The rows are filtered in order of decreasing degree, so each
    H apply covers a shrinking prefix of the block.  On 2-16 atom cells this
    saves 16-18% of the band applies at the same accuracy of the k.p masses.
    buffer_degree: "cap" filters the buffer rows with `degree`, "wanted" with
    the largest degree an unconverged wanted row needs this step (5-15%
    fewer band applies when every wanted row has to converge)."""
    dev = X0.device
    cdt = torch.complex128
    wdt = work_dtype or X0.dtype
    nb = X0.shape[0]
    nconv = nb if nconv is None else nconv
    gr = gram or _gram
    eig = eigh or (lambda A: torch.linalg.eigh(0.5 * (A + A.conj().T)))
    lc = lincomb or (lambda z, a, y, b, x, c: z.mul_(a).add_(y, alpha=b).add_(x, alpha=c))

    def Hw(V):                               # work precision in and out
        HV = torch.empty_like(V)
        apply(V, HV)
        return HV

    def H(V):
        return Hw(V.to(wdt).contiguous()).to(cdt)

    b_up = _spectral_upper_bound(H, X0.shape[1], dev)
    X, _ = _chol_orth(X0.to(cdt), gram=gram)
    HX = H(X)
    e, C = eig(gr(X, HX))
    X = C.T @ X
    HX = C.T @ HX
    it = 0
    for it in range(1, maxiter + 1):
        R = HX - e[:, None] * X
        rn = torch.linalg.norm(R, dim=1)
        if int((rn[:nconv] < tol).sum()) == nconv:
            break
        a, a0 = float(e[-1]), float(e[0])
        b = max(b_up, a + 1e-3 * max(1.0, abs(a)))
        half, mid = 0.5 * (b - a), 0.5 * (b + a)
        sigma = half / (a0 - mid)
        tau = 2.0 / sigma
        if adaptive:
            t = ((e - mid) / half).abs().clamp(min=1.0 + 1e-12)
            need = torch.ceil(torch.acosh((rn / tol).clamp(min=1.0)) / torch.acosh(t)) + 1
            m = need.clamp(min_degree, degree).long()
            if buffer_degree == "wanted":
                # the buffer rows only shape the subspace: the largest degree
                # any unconverged wanted row needs this step
                act = m[:nconv][rn[:nconv] >= tol]
                m[nconv:] = int(act.max()) if act.numel() else min_degree
            else:
                m[nconv:] = degree
            m[:nconv][rn[:nconv] < tol] = 0
            if uniform:
                m[m > 0] = int(m.max())
            order = torch.argsort(m, descending=True, stable=True)
            ms = m[order].tolist()
            if min_rows > 0:
                top = ms[min(min_rows, nb) - 1]
                if top > 0:
                    ms = [ms[0] if i < min_rows else v for i, v in enumerate(ms)]
            X, HX = X[order], HX[order]
        else:
            ms = [degree] * nb
        # rows [n_s, n_{s-1}) stop at degree s - 1 and are copied to F then;
        # rows of degree 0 keep X
        F = X.to(wdt, copy=True)
        nprev = sum(1 for v in ms if v >= 1)
        # Y1 = (H - mid) X sigma / half, from the FP64 HX of the last step
        Y = ((HX[:nprev] - mid * X[:nprev]) * (sigma / half)).to(wdt).contiguous()
        Xp = F[:nprev]
        for s in range(2, ms[0] + 1):
            n = sum(1 for v in ms if v >= s)
            if n < nprev:
                F[n:nprev] = Y[n:nprev]
            sig_new = 1.0 / (tau - sigma)
            # Y_next = (H - mid) Y 2 sig_new / half - sigma sig_new Y_prev, in place
            Yn = Hw(Y[:n])
            lc(Yn, 2.0 * sig_new / half, Y[:n], -mid * 2.0 * sig_new / half, Xp[:n], -sigma * sig_new)
            Xp, Y, sigma, nprev = Y, Yn, sig_new, n
        F[:nprev] = Y[:nprev]
        Y = F.to(cdt)
        try:
            X, _ = _chol_orth(Y, gram=gram)
        except RuntimeError:
            X, _ = _chol_orth(torch.linalg.qr(Y.T)[0].T.contiguous(), gram=gram)
        HX = H(X)
        e, C = eig(gr(X, HX))
        X = C.T @ X
        HX = C.T @ HX
    R = HX - e[:, None] * X
    rn = torch.linalg.norm(R, dim=1)
    return EigResult(E=e, X=X, HX=HX, resid=rn, iters=it, converged=int((rn[:nconv] < tol).sum()))


def plane_wave_start(apply, kin: torch.Tensor, rows: int, m: int | None = None, work_dtype=torch.complex64,
                     eigh=None) -> torch.Tensor:
    """Start vectors from a Rayleigh-Ritz step in the m (default 1.25 rows)
    plane waves of lowest kinetic energy: H is applied to those unit
    vectors (one block of m band applies in the work precision), the m x m
    matrix <G|H|G'> is diagonalised in FP64 and its `rows` lowest
    eigenvectors, placed on those plane waves, start the iteration
    (complex128, rows x npw).  The low-lying states of a smooth potential
    live mostly on the low-|k+G| plane waves, so the filtered iteration
    starts near the wanted subspace: on the 8-64 atom serving cells it
    needs 25-45% fewer band applies than a random start.  (m = rows took
    7 instead of 4 filter steps on the 64-atom cell; 1.25 rows needs as few
    band applies as 2 rows, at a host eigensolve half as long.)"""
    npw = int(kin.numel())
    m = min(npw, max(rows, m if m is not None else rows + rows // 4))
    idx = torch.argsort(kin.to(torch.float64), stable=True)[:m]
    ar = torch.arange(m, device=kin.device)
    E = torch.zeros(m, npw, dtype=work_dtype, device=kin.device)
    E[ar, idx] = 1.0
    HE = torch.empty_like(E)
    apply(E, HE)
    # row i of HE is H e_idx[i], so HE[i, idx[j]] = <G_j|H|G_i>
    Hs = HE[:, idx].T.to(torch.complex128)
    _, c = (eigh or torch.linalg.eigh)(0.5 * (Hs + Hs.conj().T))
    X0 = torch.zeros(rows, npw, dtype=torch.complex128, device=kin.device)
    X0[:, idx] = c[:, :rows].T.to(X0.device)
    return X0


def random_start(nb, npw, kin, seed=0, device="cpu"):
    """Deterministic start vectors: complex normal weighted by 1/(1+T(G)),
    drawn on `device` (the same seed gives the same vectors on the same
    kind of device)."""
    dev = torch.device(device)
    g = torch.Generator(device=dev).manual_seed(seed)
    re = torch.randn(nb, npw, generator=g, dtype=torch.float64, device=dev)
    im = torch.randn(nb, npw, generator=g, dtype=torch.float64, device=dev)
    w = 1.0 / (1.0 + kin.detach().to(device=dev, dtype=torch.float64))
    return torch.complex(re, im) * w[None, :]
