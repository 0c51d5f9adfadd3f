"""Seeded synthetic inputs for the H|psi> path.

The workloads (cells, cutoffs, potentials, projectors, wavefunctions) are
generated deterministically from (name, seed) with a SplitMix64 stream and
Box-Muller normals, so the device path and the CPU oracle see identical bytes.

Conventions follow the reference:
  * K = k + G in Cartesian bohr^-1, kinetic = |K|^2 / 2 Hartree
    (src/io/GKK.cppm:543-549, 652-671);
  * centred integer G components, grid index ((g mod n) + n) mod n with the
    k axis fastest (src/H_psi/hamiltonian_callable.cpp:141-147);
  * V_loc(r) in Hartree on [i][j][k] (src/io/RHO.cppm:57-61);
  * beta(q) tables on q_i = i*dq, dq = 0.01, q_max = 20, nq = ceil(q_max/dq)
    + 4 nodes, raw values without 1/sqrt(Omega)
    (src/pseudo/ncpp_advance.cppm:586-601);
  * D is used only through its per-l blocks, diagonalised and folded into
    rotated beta tables (src/pseudo/ncpp_advance.cppm:65-153).

Energies in the workload descriptions are in Rydberg and converted to Hartree
here (1 Ha = 2 Ry).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

RY_PER_HA = 2.0
DQ = 0.01
QMAX = 20.0
NQ = int(math.ceil(QMAX / DQ)) + 1 + 3   # 2004


# --------------------------------------------------------------------------
#  SplitMix64 stream
# --------------------------------------------------------------------------
class SplitMix64:
    """Counter-based SplitMix64: draw i of the stream is mix(seed + i*gamma)."""

    GAMMA = np.uint64(0x9E3779B97F4A7C15)

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            i = np.arange(self.counter + 1, self.counter + n + 1, dtype=np.uint64)
            z = self.seed + i * self.GAMMA
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.counter += n
        return z

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0, 1)."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)          # (0, 1]
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


# --------------------------------------------------------------------------
#  Data containers (host side, numpy)
# --------------------------------------------------------------------------
@dataclass
class Species:
    z: int
    lmax: int
    beta_l: np.ndarray          # int32 [nbeta], sorted by l
    lam: np.ndarray             # float64 [nbeta] Hartree (diagonalised D)
    beta_tab: np.ndarray        # float64 [nbeta, nq]
    dbeta_tab: np.ndarray       # float64 [nbeta, nq]
    dq: float = DQ
    qmax: float = QMAX
    # closed form of the tables: tab[k](q) = sum_i coef[k, i] * c * q^ls[i] * exp(-w[i] q^2)
    raw: dict | None = None

    def beta_exact(self, k: int, q: np.ndarray, deriv: bool = False) -> np.ndarray:
        """Closed-form (rotated) beta_k(q) or dbeta_k/dq, without 1/sqrt(Omega)."""
        r = self.raw
        q = np.asarray(q, dtype=np.float64)
        out = np.zeros_like(q)
        for i, (l, w) in enumerate(zip(r["ls"], r["w"])):
            e = np.exp(-w * q * q)
            if deriv:
                f = ((l * q ** (l - 1)) if l > 0 else 0.0) - 2.0 * w * q ** (l + 1)
            else:
                f = q ** l
            out = out + r["coef"][k, i] * r["c"] * f * e
        return out

    @property
    def nbeta(self) -> int:
        return int(self.beta_l.shape[0])

    @property
    def nq(self) -> int:
        return int(self.beta_tab.shape[1])

    def nchannels(self) -> int:
        """Projector channels per atom, skipping |lambda| < 1e-15 as the
        reference does (src/H_psi/hamiltonian_callable.cpp:227)."""
        return int(sum(2 * l + 1 for l, lam in zip(self.beta_l, self.lam) if abs(lam) >= 1e-15))


@dataclass
class Cell:
    A: np.ndarray               # [3,3] rows = lattice vectors (bohr)
    n: tuple                    # FFT grid (n1, n2, n3)
    vloc: np.ndarray            # float64 [n1*n2*n3] Hartree
    tau: np.ndarray             # float64 [natom, 3] fractional
    species: np.ndarray         # int32 [natom] index into species_list
    species_list: list = field(default_factory=list)

    @property
    def omega(self) -> float:
        return float(abs(np.linalg.det(self.A)))

    @property
    def B(self) -> np.ndarray:
        """Reciprocal lattice, rows b_i with a_i . b_j = 2 pi delta_ij."""
        return 2.0 * np.pi * np.linalg.inv(self.A).T

    @property
    def natom(self) -> int:
        return int(self.tau.shape[0])

    def nchannels(self) -> int:
        return int(sum(self.species_list[s].nchannels() for s in self.species))


@dataclass
class KPoint:
    kfrac: np.ndarray           # [3]
    K: np.ndarray               # float64 [npw, 3]
    kin: np.ndarray             # float64 [npw]
    gidx: np.ndarray            # int32 [npw, 3] centred

    @property
    def npw(self) -> int:
        return int(self.K.shape[0])


@dataclass
class Workload:
    name: str
    cell: Cell
    kpts: list
    nb: int
    ecut_ry: float
    seed: int
    psi: list = field(default_factory=list)     # per k: complex128 [nb, npw]


# --------------------------------------------------------------------------
#  Cells
# --------------------------------------------------------------------------
DIAMOND8 = np.array([
    [0.00, 0.00, 0.00], [0.00, 0.50, 0.50], [0.50, 0.00, 0.50], [0.50, 0.50, 0.00],
    [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25],
])


def fcc_primitive(a: float):
    A = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    tau = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    return A, tau


def diamond_cubic(a: float, rep: int = 1):
    """rep x rep x rep supercell of the 8-atom diamond cell of edge a."""
    A = a * rep * np.eye(3)
    pos = []
    for i in range(rep):
        for j in range(rep):
            for k in range(rep):
                for f in DIAMOND8:
                    pos.append((f + np.array([i, j, k])) / rep)
    return A, np.array(pos)


def g_sphere(A: np.ndarray, n, kfrac, ecut_ry: float, order: str = "kinetic"):
    """All centred integer G in the FFT box with |k+G|^2 <= ecut (Ry units:
    |K|^2 in bohr^-2 equals the kinetic energy in Ry).  order "kinetic":
    sorted by |K|^2 then by index (stable), the usual PWmat/QE sphere order;
    "plane": by FFT-grid plane (g1 mod n1), row (g2 mod n2) and column
    (g3 mod n3), so each plane's and row's plane waves are consecutive (the
    device scatter / gather then reads and writes contiguous runs)."""
    B = 2.0 * np.pi * np.linalg.inv(A).T
    axes = [np.arange(-(m // 2), m - m // 2) for m in n]
    g1, g2, g3 = np.meshgrid(*axes, indexing="ij")
    g = np.stack([g1.ravel(), g2.ravel(), g3.ravel()], axis=1)
    K = (g + np.asarray(kfrac)[None, :]) @ B
    k2 = np.einsum("ij,ij->i", K, K)
    keep = np.nonzero(k2 <= ecut_ry * (1.0 + 1e-14))[0]
    if order == "kinetic":
        sel = keep[np.argsort(k2[keep], kind="stable")]
    elif order == "plane":
        gm = np.mod(g[keep], np.asarray(n)[None, :])
        sel = keep[np.argsort((gm[:, 0] * n[1] + gm[:, 1]) * n[2] + gm[:, 2], kind="stable")]
    else:
        raise ValueError(f"unknown sphere order {order!r}")
    K = K[sel]
    return KPoint(kfrac=np.asarray(kfrac, dtype=np.float64),
                  K=np.ascontiguousarray(K),
                  kin=0.5 * np.einsum("ij,ij->i", K, K),
                  gidx=np.ascontiguousarray(g[sel].astype(np.int32)))


def make_vloc(A, n, tau, rng: SplitMix64, amp_ry=-2.0, sigma=1.0,
              noise_ry=0.05, gcut_noise=4.0):
    """Gaussians of width sigma at every atom (all periodic images) plus
    N(0, noise) white noise low-passed to |G| < gcut_noise; real; Hartree."""
    n1, n2, n3 = n
    omega = abs(np.linalg.det(A))
    B = 2.0 * np.pi * np.linalg.inv(A).T
    f = [np.fft.fftfreq(m, 1.0 / m).astype(np.int64) for m in n]   # FFT order ints
    g1, g2, g3 = np.meshgrid(*f, indexing="ij")
    G = np.stack([g1, g2, g3], axis=-1) @ B
    G2 = np.einsum("...i,...i->...", G, G)
    # structure factor sum_t p1_t(i) p2_t(j) p3_t(k) as one matrix product
    # over the atoms: [(i, j), t] x [t, k]
    tau = np.asarray(tau, dtype=np.float64).reshape(-1, 3)
    p1, p2, p3 = (np.exp(-2j * np.pi * np.outer(tau[:, a], f[a])) for a in range(3))
    sf = ((p1[:, :, None] * p2[:, None, :]).reshape(len(tau), n1 * n2).T @ p3).reshape(n)
    vg = (amp_ry / RY_PER_HA) * (2 * np.pi * sigma ** 2) ** 1.5 / omega * np.exp(-0.5 * sigma ** 2 * G2) * sf
    noise = rng.normal(n1 * n2 * n3).reshape(n) * (noise_ry / RY_PER_HA)
    nk = np.fft.fftn(noise)
    nk[G2 >= gcut_noise ** 2] = 0.0
    # the Gaussians' and the filtered noise's spectra back in one transform
    v = np.fft.ifftn(vg * (n1 * n2 * n3) + nk).real
    return np.ascontiguousarray(v.ravel())


def make_species(rng: SplitMix64, z=14, ls=(0, 0, 1, 1), widths=(0.25, 0.5), c=1.0):
    """Projectors beta_{l,ib}(q) = c q^l exp(-w_ib q^2) with a random
    symmetric dion (entries U(-1,1) Ry) over all channels; as in the
    reference only the per-l blocks survive, are diagonalised, and rotate the
    beta tables (src/pseudo/ncpp_advance.cppm:65-153)."""
    ls = np.asarray(ls, dtype=np.int32)
    nb = len(ls)
    u = rng.uniform(nb * nb).reshape(nb, nb) * 2.0 - 1.0
    dion = 0.5 * (u + u.T) / RY_PER_HA
    q = np.arange(NQ) * DQ
    raw, draw = [], []
    ib_count = {}
    for l in ls:
        ib = ib_count.get(int(l), 0)
        ib_count[int(l)] = ib + 1
        w = widths[ib % len(widths)]
        e = np.exp(-w * q * q)
        raw.append(c * q ** l * e)
        dl = (l * q ** (l - 1) if l > 0 else np.zeros_like(q)) - 2.0 * w * q ** (l + 1)
        draw.append(c * dl * e)
    raw = np.array(raw)
    draw = np.array(draw)
    lam = np.zeros(nb)
    tab = np.zeros_like(raw)
    dtab = np.zeros_like(draw)
    coef = np.zeros((nb, nb))
    wl = []
    cnt = {}
    for l in ls:
        ib = cnt.get(int(l), 0)
        cnt[int(l)] = ib + 1
        wl.append(widths[ib % len(widths)])
    for l in sorted(set(int(x) for x in ls)):
        idx = np.nonzero(ls == l)[0]
        w, U = np.linalg.eigh(dion[np.ix_(idx, idx)])
        for k, ik in enumerate(idx):
            lam[ik] = w[k]
            tab[ik] = U[:, k] @ raw[idx]
            dtab[ik] = U[:, k] @ draw[idx]
            coef[ik, idx] = U[:, k]
    return Species(z=z, lmax=int(ls.max()), beta_l=ls, lam=lam,
                   beta_tab=np.ascontiguousarray(tab), dbeta_tab=np.ascontiguousarray(dtab),
                   raw={"ls": [int(x) for x in ls], "w": wl, "c": c, "coef": coef, "dion": dion})


def make_psi(kp: KPoint, nb: int, rng: SplitMix64) -> np.ndarray:
    """complex N(0,1) in the sphere, weighted by 1/(1+|K|^2), orthonormalised
    (sum_G conj(psi_m) psi_n = delta_mn)."""
    npw = kp.npw
    w = 1.0 / (1.0 + 2.0 * kp.kin)
    x = rng.normal(2 * nb * npw).reshape(nb, npw, 2)
    z = (x[..., 0] + 1j * x[..., 1]) * w[None, :]
    if nb <= npw:
        q, _ = np.linalg.qr(z.T)
        z = q.T
    return np.ascontiguousarray(z.astype(np.complex128))


# --------------------------------------------------------------------------
#  The five workloads
# --------------------------------------------------------------------------
# The named workloads: cell, grid, cutoff, bands, k-points.
#   si2_fcc     2-atom fcc Si, a = 10.26 bohr, 30 Ry, 36^3, 8 bands, Gamma
#   si64_cubic  64-atom diamond supercell, a = 20.52 bohr, 40 Ry, 90^3, 256
#               bands, k = (0.1, 0, 0)
#   si64_kgrid  the same cell, 4x4x2 k-grid (32 k-points), 128 bands
#   si8_cubic   8-atom cubic Si, a = 10.26 bohr, 50 Ry, 48^3, 200 bands, Gamma
WORKLOAD_SEEDS = {"si2_fcc": 0xEFF0, "si64_cubic": 0xEFF1, "si64_kgrid": 0xEFF2, "si8_cubic": 0xEFF3}


def mp_grid(m1, m2, m3):
    ks = []
    for i in range(m1):
        for j in range(m2):
            for k in range(m3):
                f = np.array([i / m1, j / m2, k / m3])
                f = np.where(f > 0.5 + 1e-12, f - 1.0, f)
                ks.append(f)
    return ks


def build(name: str, nb: int | None = None, nk: int | None = None, seed: int | None = None,
          with_psi: bool = True, two_species: bool = False, order: str = "kinetic") -> Workload:
    """Build the named workload (WORKLOAD_SEEDS).  nb / nk shrink the band
    or k-point count for quick parity cases; the cell, grid and cutoff never
    change; `order` is the sphere order of g_sphere."""
    if name not in WORKLOAD_SEEDS:
        raise ValueError(f"unknown workload {name!r}")
    seed = WORKLOAD_SEEDS[name] if seed is None else seed
    rng = SplitMix64(seed)
    if name == "si2_fcc":
        A, tau = fcc_primitive(10.26)
        n, ecut, nbd, ks = (36, 36, 36), 30.0, 8, [np.zeros(3)]
    elif name == "si64_cubic":
        A, tau = diamond_cubic(10.26, rep=2)
        n, ecut, nbd, ks = (90, 90, 90), 40.0, 256, [np.array([0.1, 0.0, 0.0])]
    elif name == "si64_kgrid":
        A, tau = diamond_cubic(10.26, rep=2)
        n, ecut, nbd, ks = (90, 90, 90), 40.0, 128, mp_grid(4, 4, 2)
    else:
        A, tau = diamond_cubic(10.26, rep=1)
        n, ecut, nbd, ks = (48, 48, 48), 50.0, 200, [np.zeros(3)]
    if nk is not None:
        ks = ks[:nk]
    nbands = nbd if nb is None else nb
    sp0 = make_species(rng, z=14)
    species_list = [sp0]
    species = np.zeros(len(tau), dtype=np.int32)
    if two_species:
        species_list.append(make_species(rng, z=7, ls=(0, 1, 1), widths=(0.3, 0.6)))
        species[len(tau) // 2:] = 1
    vloc = make_vloc(A, n, tau, rng)
    cell = Cell(A=A, n=n, vloc=vloc, tau=np.ascontiguousarray(tau, dtype=np.float64),
                species=species, species_list=species_list)
    kpts = [g_sphere(A, n, k, ecut, order) for k in ks]
    wl = Workload(name=name, cell=cell, kpts=kpts, nb=nbands, ecut_ry=ecut, seed=seed)
    if with_psi:
        wl.psi = [make_psi(kp, nbands, rng) for kp in kpts]
    return wl


def custom(A, tau, n, ecut_ry, kfracs, nb, seed, species_idx=None, species_list=None, with_psi=True,
           order="kinetic"):
    """A workload for an arbitrary cell (used by the serving loop); `order`
    is the sphere order of g_sphere."""
    rng = SplitMix64(seed)
    if species_list is None:
        species_list = [make_species(rng)]
    species = np.zeros(len(tau), dtype=np.int32) if species_idx is None else np.asarray(species_idx, np.int32)
    vloc = make_vloc(A, n, tau, rng)
    cell = Cell(A=np.asarray(A, float), n=tuple(n), vloc=vloc,
                tau=np.ascontiguousarray(tau, dtype=np.float64), species=species,
                species_list=species_list)
    kpts = [g_sphere(cell.A, n, k, ecut_ry, order) for k in kfracs]
    wl = Workload(name="custom", cell=cell, kpts=kpts, nb=nb, ecut_ry=ecut_ry, seed=seed)
    if with_psi:
        wl.psi = [make_psi(kp, nb, rng) for kp in kpts]
    return wl
