"""Independent restatements from the mathematics (not from the reference's
code) used to pin the oracle: dense H_GG' for small cells, projectors from
closed-form beta, scipy spherical harmonics and direct exponentials."""
import numpy as np
from scipy import special


def real_ylm(l, m, theta, phi):
    """Real Y_lm with the Condon-Shortley phase: m>0 -> sqrt2 Re Y_l^m,
    m<0 -> sqrt2 Im Y_l^|m| (scipy's complex Y_l^m carries (-1)^m)."""
    y = special.sph_harm_y(l, abs(m), theta, phi)
    if m == 0:
        return y.real
    return np.sqrt(2.0) * (y.real if m > 0 else y.imag)


def spherical(K):
    q = np.linalg.norm(K, axis=1)
    th = np.where(q > 0, np.arccos(np.clip(K[:, 2] / np.where(q > 0, q, 1), -1, 1)), 0.0)
    ph = np.arctan2(K[:, 1], K[:, 0])
    return q, th, ph


def projectors(cell, kp, deriv=False):
    """Rows p_c(G) = beta(q)/sqrt(Omega) Y_lm S_a(G) in the reference channel
    order (species -> atom -> (l, ib) with |lambda| >= 1e-15 -> m), and lambda_c.
    With deriv=True also returns d p_c / d K_alpha by central differences of
    the closed-form radial*angular factor (S is k-independent in matrix
    elements)."""
    q, th, ph = spherical(kp.K)
    norm = 1.0 / np.sqrt(cell.omega)
    rows, lams, drows = [], [], []
    for s, sp in enumerate(cell.species_list):
        for a in range(cell.natom):
            if cell.species[a] != s:
                continue
            S = np.exp(-2j * np.pi * ((kp.gidx + kp.kfrac[None, :]) @ cell.tau[a]))
            for ib in range(sp.nbeta):
                lam = sp.lam[ib]
                if abs(lam) < 1e-15:
                    continue
                l = int(sp.beta_l[ib])
                for m in range(-l, l + 1):
                    f = sp.beta_exact(ib, q) * norm * real_ylm(l, m, th, ph)
                    rows.append(f * S)
                    lams.append(lam)
                    if deriv:
                        d = []
                        h = 1e-5
                        for al in range(3):
                            Kp = kp.K.copy(); Kp[:, al] += h
                            Km = kp.K.copy(); Km[:, al] -= h
                            qp, tp, pp = spherical(Kp)
                            qm, tm, pm = spherical(Km)
                            fp = sp.beta_exact(ib, qp) * norm * real_ylm(l, m, tp, pp)
                            fm = sp.beta_exact(ib, qm) * norm * real_ylm(l, m, tm, pm)
                            d.append((fp - fm) / (2 * h) * S)
                        drows.append(d)
    P = np.array(rows) if rows else np.zeros((0, kp.npw), complex)
    if deriv:
        return P, np.array(lams), np.array(drows)
    return P, np.array(lams)


def vloc_matrix(cell, kp):
    """V(G-G') with V~(G) = (1/N) sum_r V(r) exp(-i G.r) (the convolution
    identity that fixes the FFT sign and normalisation)."""
    n1, n2, n3 = cell.n
    vg = np.fft.fftn(cell.vloc.reshape(cell.n)) / (n1 * n2 * n3)
    d = kp.gidx[:, None, :] - kp.gidx[None, :, :]
    return vg[d[..., 0] % n1, d[..., 1] % n2, d[..., 2] % n3]


def dense_h(cell, kp, local=True, nonlocal_=True):
    H = np.zeros((kp.npw, kp.npw), complex)
    if local:
        H += np.diag(kp.kin) + vloc_matrix(cell, kp)
    if nonlocal_:
        P, lam = projectors(cell, kp)
        if len(lam):
            H += (P.T * lam) @ P.conj()
    return H


def dense_vnl_raw(cell, kp):
    """V_NL = sum_atoms sum_{i,j: l_i = l_j} sum_m |p_im> D_ij <p_jm| from the
    UNROTATED closed-form radial functions c q^l exp(-w q^2) of each raw
    channel and the raw dion of make_species (D in Hartree), with the
    cross-l entries of D dropped, as the reference's diagonalizeNonlocal
    does (src/pseudo/ncpp_advance.cppm:65-153).  Independent of the
    eigen-decomposition the engine consumes (lambda, rotated tables)."""
    q, th, ph = spherical(kp.K)
    norm = 1.0 / np.sqrt(cell.omega)
    V = np.zeros((kp.npw, kp.npw), complex)
    for s, sp in enumerate(cell.species_list):
        r = sp.raw
        ls, ws, c, D = r["ls"], r["w"], r["c"], r["dion"]
        nb = len(ls)
        for a in range(cell.natom):
            if cell.species[a] != s:
                continue
            S = np.exp(-2j * np.pi * ((kp.gidx + kp.kfrac[None, :]) @ cell.tau[a]))
            for i in range(nb):
                for j in range(nb):
                    if ls[i] != ls[j] or D[i, j] == 0.0:
                        continue
                    l = ls[i]
                    bi = c * q ** l * np.exp(-ws[i] * q * q) * norm
                    bj = c * q ** l * np.exp(-ws[j] * q * q) * norm
                    for m in range(-l, l + 1):
                        y = real_ylm(l, m, th, ph)
                        pi = bi * y * S
                        pj = bj * y * S
                        V += D[i, j] * np.outer(pi, pj.conj())
    return V
