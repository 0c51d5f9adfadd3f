"""PWmat / UPF inputs -> device-resident H|psi> data.

Readers for the files the reference loads (`Hamiltonian::loadFromDirectory`,
src/H_psi/hamiltonian.cpp:131-153) and the setup its `finalize()` does
(src/H_psi/hamiltonian.cpp:267-392), producing the `synth.Cell` /
`synth.KPoint` containers the engine uploads:

  OUT.GKK      Fortran records: header (n1,n2,n3,mg_nx,nnode,nkpt,is_SO,islda),
               Ecut, AL(3,3) in Angstrom, ngtotnod, then per k and node the
               records gkk, gkk_x, gkk_y, gkk_z holding -(k+G)
               (src/io/GKK.cppm:303-333, 511-593).
  OUT.VR/RHO   header (n1,n2,n3,nnode[,nstate]), AL, per node a slice of
               the flat [i][j][k] grid in Hartree (src/io/RHO.cppm:57-61,
               143-193).
  OUT.WG       header (n1,n2,n3,nband,mg_nx,nnode,nkpt,is_SO,islda), Ecut,
               AL, ngtotnod, then per k, node and band one record of mg_nx
               complex128 coefficients (first ngtotnod valid); with spin-
               orbit the record holds the up half then the down half and
               the header's mg_nx / ngtotnod count both (src/io/WG.cppm:
               251-277, 280-307, 320-337, 360-411).
  OUT.EIGEN    header, per spin and k a (islda, ikpt, weight, k) record and
               the eigenvalues in eV (src/io/EIGEN.cppm:164-232).
  atom.config  natom, LATTICE rows (Angstrom), POSITION Z + fractional xyz
               (src/io/ATOM.cppm:200-267, species in ascending Z, atoms in
               stable file order, :290-316).
  UPF v2       mesh r / rab, PP_BETA (r*beta(r)) with angular_momentum and
               cutoff_radius_index, PP_DIJ in Ry (src/pseudo/ncpp.cppm:163-242).

Projector setup follows sortByL + diagonalizeNonlocal
(src/pseudo/ncpp_advance.cppm:11-153: stable sort by l, per-l eigen-
decomposition of D, cross-l entries dropped, beta rotated) and the
Fourier-Bessel tables of BetaqTables / DBetaqTables
(src/pseudo/ncpp_advance.cppm:524-641, 718-835: q_i = i*dq, nq =
ceil(qmax/dq) + 4, value 4 pi int beta(r) r j_l(q r) dr by Simpson's rule
over the UPF mesh).
"""
from __future__ import annotations

import math
import struct
import xml.etree.ElementTree as ET
from dataclasses import dataclass

import numpy as np
from scipy import special

from . import synth

BOHR_RADIUS_ANGSTROM = 0.52917721067
HARTREE_TO_EV = 27.211386245988
RY_PER_HA = 2.0


# --------------------------------------------------------------------------
#  Fortran unformatted records (4-byte length markers)
# --------------------------------------------------------------------------
class _Records:
    def __init__(self, path):
        self.buf = open(path, "rb").read()
        self.pos = 0

    def next(self) -> bytes:
        b, o = self.buf, self.pos
        if o + 4 > len(b):
            raise RuntimeError("Failed to read record length marker")
        (n,) = struct.unpack_from("<i", b, o)
        end = o + 4 + n
        if end + 4 > len(b):
            raise RuntimeError("record extends past end of file")
        (n2,) = struct.unpack_from("<i", b, end)
        if n2 != n:
            raise RuntimeError("Record length mismatch")
        self.pos = end + 4
        return b[o + 4:end]


def _lattice_from_al(al_bytes) -> np.ndarray:
    """AL(3,3) column-major in Angstrom -> rows a_n in bohr (lattice.cppm:113-122)."""
    flat = np.frombuffer(al_bytes, dtype="<f8", count=9)
    return flat.reshape(3, 3) / BOHR_RADIUS_ANGSTROM


# --------------------------------------------------------------------------
#  OUT.GKK
# --------------------------------------------------------------------------
@dataclass
class GKKMeta:
    n: tuple
    mg_nx: int
    nnode: int
    nkpt: int
    is_so: int
    islda: int
    ecut: float
    A: np.ndarray           # bohr, rows


class GKKFile:
    def __init__(self, path):
        rec = _Records(path)
        h = struct.unpack("<8i", rec.next())
        n1, n2, n3, mg_nx, nnode, nkpt, is_so, islda = h
        if is_so == 1:
            mg_nx //= 2
        (ecut,) = struct.unpack("<d", rec.next())
        A = _lattice_from_al(rec.next())
        ng = np.frombuffer(rec.next(), dtype="<i4")
        if ng[0] != nnode:
            raise RuntimeError("nnodes mismatch")
        ngt = ng[1:1 + nkpt * nnode].reshape(nkpt, nnode).astype(np.int64)
        if is_so == 1:
            ngt //= 2
        self.meta = GKKMeta(n=(n1, n2, n3), mg_nx=mg_nx, nnode=nnode, nkpt=nkpt, is_so=is_so,
                            islda=islda, ecut=ecut, A=A)
        self.ngtotnod = ngt
        self._kdata = []
        for ik in range(nkpt):
            parts = [[], [], [], []]
            for inode in range(nnode):
                m = int(ngt[ik, inode])
                for a in range(4):
                    r = rec.next()
                    if m:
                        if 8 * m > len(r):
                            raise RuntimeError("gkk: nbytes exceeds record length")
                        parts[a].append(np.frombuffer(r, dtype="<f8", count=m))
            self._kdata.append([np.concatenate(p) if p else np.zeros(0) for p in parts])

    @property
    def nkpt(self):
        return self.meta.nkpt

    def kpoint(self, ikpt: int) -> synth.KPoint:
        """GKK::loadKPoint with the Cartesian, Spherical and Integer views."""
        if ikpt < 0 or ikpt >= self.nkpt:
            raise IndexError(f"Invalid k-point index: {ikpt}")
        kin, gx, gy, gz = self._kdata[ikpt]
        K = -np.stack([gx, gy, gz], axis=1)           # stored -(k+G)  (GKK.cppm:543-549)
        if K.shape[0] == 0:
            raise RuntimeError("current k-point has no G-vectors")
        kf = infer_kfrac(K, self.meta.A)
        gidx = integer_indices(K, self.meta.A, kf)
        return synth.KPoint(kfrac=kf, K=np.ascontiguousarray(K), kin=kin.copy(), gidx=gidx)


def infer_kfrac(K: np.ndarray, A: np.ndarray) -> np.ndarray:
    """Fractional k from the circular mean of frac(A.K / 2 pi), wrapped to
    (-0.5, 0.5] with a 1e-12 tolerance at 0.5 (GKK.cppm:596-649)."""
    c = (K @ A.T) / (2.0 * np.pi)
    d = c - np.floor(c)
    th = d * 2.0 * np.pi
    avg = np.arctan2(np.sin(th).sum(axis=0), np.cos(th).sum(axis=0))
    avg = np.where(avg < 0.0, avg + 2.0 * np.pi, avg)
    k = avg / (2.0 * np.pi)
    return np.where(k > 0.5 + 1e-12, k - 1.0, k)


def integer_indices(K, A, kfrac) -> np.ndarray:
    """g = lround(A.K / 2 pi - k_frac) (GKK.cppm:452-473); lround rounds
    halves away from zero."""
    c = (K @ A.T) / (2.0 * np.pi) - kfrac[None, :]
    return (np.sign(c) * np.floor(np.abs(c) + 0.5)).astype(np.int32)


# --------------------------------------------------------------------------
#  OUT.WG
# --------------------------------------------------------------------------
@dataclass
class WGMeta:
    n: tuple
    nband: int
    mg_nx: int              # per spin half
    nnode: int
    nkpt: int
    is_so: int
    islda: int
    ecut: float
    A: np.ndarray           # bohr, rows
    ng_per_kpt: np.ndarray  # coefficients per k-point (per spin half)


class WGFile:
    """OUT.WG, memory-mapped: the record offsets of every (k, node, band) are
    indexed once (WG::computeOffsets, src/io/WG.cppm:320-337); bands are
    read on demand.  Coefficients are in the order of the matching OUT.GKK
    k-point (nodes concatenated)."""

    def __init__(self, path):
        self._mm = np.memmap(path, dtype=np.uint8, mode="r")
        pos = 0

        def rec(p):
            if p + 4 > self._mm.size:
                raise RuntimeError("Failed to read record length marker")
            (n,) = struct.unpack_from("<i", self._mm, p)
            end = p + 4 + n
            if n < 0 or end + 4 > self._mm.size:
                raise RuntimeError("record extends past end of file")
            (n2,) = struct.unpack_from("<i", self._mm, end)
            if n2 != n:
                raise RuntimeError("Record length mismatch")
            return p + 4, n, end + 4

        d, n, pos = rec(pos)
        if n < 36:
            raise RuntimeError("header: read failed")
        n1, n2, n3, nband, mg_nx, nnode, nkpt, is_so, islda = struct.unpack_from("<9i", self._mm, d)
        if is_so == 1:
            mg_nx //= 2
        d, n, pos = rec(pos)
        (ecut,) = struct.unpack_from("<d", self._mm, d)
        d, n, pos = rec(pos)
        A = _lattice_from_al(bytes(self._mm[d:d + 72]))
        d, n, pos = rec(pos)
        ng = np.frombuffer(bytes(self._mm[d:d + n]), dtype="<i4")
        if ng[0] != nnode:
            raise RuntimeError("nnodes mismatch")
        ngt = ng[1:1 + nkpt * nnode].reshape(nkpt, nnode).astype(np.int64)
        if is_so == 1:
            ngt //= 2
        self.meta = WGMeta(n=(n1, n2, n3), nband=nband, mg_nx=mg_nx, nnode=nnode, nkpt=nkpt, is_so=is_so,
                           islda=islda, ecut=ecut, A=A, ng_per_kpt=ngt.sum(axis=1))
        self.ngtotnod = ngt
        # file layout per k: node 0 bands 0..nband-1, node 1 bands ..., ...
        self._off = np.zeros((nkpt, nnode, nband), dtype=np.int64)
        self._len = np.zeros((nkpt, nnode, nband), dtype=np.int64)
        for ik in range(nkpt):
            for inode in range(nnode):
                for ib in range(nband):
                    d, n, pos = rec(pos)
                    self._off[ik, inode, ib] = d
                    self._len[ik, inode, ib] = n

    def band(self, ikpt: int, iband: int):
        """(up, down) coefficients of one band; down is None without
        spin-orbit (WG::loadBand, src/io/WG.cppm:360-411)."""
        m = self.meta
        if ikpt < 0 or ikpt >= m.nkpt:
            raise IndexError(f"Invalid k-point index: {ikpt}")
        if iband < 0 or iband >= m.nband:
            raise IndexError(f"Invalid band index: {iband}")
        up, down = [], []
        for inode in range(m.nnode):
            g = int(self.ngtotnod[ikpt, inode])
            if g == 0:
                continue
            o, n = int(self._off[ikpt, inode, iband]), int(self._len[ikpt, inode, iband])
            if m.is_so == 1:
                half = n // 32
                if g > half:
                    raise RuntimeError("wg: nbytes exceeds record length")
                up.append(np.frombuffer(self._mm, dtype="<c16", count=g, offset=o))
                down.append(np.frombuffer(self._mm, dtype="<c16", count=g, offset=o + 16 * half))
            else:
                if 16 * g > n:
                    raise RuntimeError("wg: nbytes exceeds record length")
                up.append(np.frombuffer(self._mm, dtype="<c16", count=g, offset=o))
        u = np.concatenate(up) if up else np.zeros(0, np.complex128)
        if m.is_so != 1:
            return u.copy(), None
        return u.copy(), np.concatenate(down) if down else np.zeros(0, np.complex128)

    def bands(self, ikpt: int, which=None, spin: str = "up") -> np.ndarray:
        """psi [nb, ng] (band-major, the layout effm_hpsi takes) for the
        bands `which` (default all) of k-point ikpt."""
        which = range(self.meta.nband) if which is None else which
        k = 0 if spin == "up" else 1
        rows = [self.band(ikpt, ib)[k] for ib in which]
        if any(r is None for r in rows):
            raise ValueError("no spin-down half without spin-orbit")
        return np.ascontiguousarray(np.stack(rows)) if rows else np.zeros((0, 0), np.complex128)


def write_wg(path, n, A_bohr, ecut, coeffs, node_split=None, mg_nx=None, islda=1):
    """Write an OUT.WG in the layout WGFile reads.  coeffs[k] is [nband, ng]
    (or an (up, down) pair of them for spin-orbit); node_split[k] lists the
    coefficient counts per node (default one node).  Records hold mg_nx
    (per half) complex128, zero-padded past each node's count."""
    so = isinstance(coeffs[0], tuple)
    nk = len(coeffs)
    up0 = coeffs[0][0] if so else coeffs[0]
    nband = up0.shape[0]
    splits = node_split or [[(c[0] if so else c).shape[1]] for c in coeffs]
    nnode = len(splits[0])
    mg = mg_nx or max(max(s) for s in splits)

    def rec(f, payload: bytes):
        f.write(struct.pack("<i", len(payload)))
        f.write(payload)
        f.write(struct.pack("<i", len(payload)))

    fac = 2 if so else 1
    with open(path, "wb") as f:
        rec(f, struct.pack("<9i", *n, nband, fac * mg, nnode, nk, 1 if so else 0, islda))
        rec(f, struct.pack("<d", ecut))
        rec(f, (np.asarray(A_bohr, float) * BOHR_RADIUS_ANGSTROM).astype("<f8").tobytes())
        ngt = np.array([[fac * g for g in s] for s in splits], dtype="<i4")
        rec(f, struct.pack("<i", nnode) + ngt.tobytes())
        for ik in range(nk):
            up = coeffs[ik][0] if so else coeffs[ik]
            dn = coeffs[ik][1] if so else None
            starts = np.concatenate([[0], np.cumsum(splits[ik])])
            for inode in range(nnode):
                a, b = starts[inode], starts[inode + 1]
                for ib in range(nband):
                    buf = np.zeros(fac * mg, dtype="<c16")
                    buf[:b - a] = up[ib, a:b]
                    if so:
                        buf[mg:mg + b - a] = dn[ib, a:b]
                    rec(f, buf.tobytes())


# --------------------------------------------------------------------------
#  OUT.VR / OUT.RHO
# --------------------------------------------------------------------------
def read_vr(path):
    """Return (n, A_bohr, data[nstate, n1*n2*n3]) in file order."""
    rec = _Records(path)
    h = rec.next()
    if len(h) == 20:
        n1, n2, n3, nnode, nstate = struct.unpack("<5i", h)
    elif len(h) == 16:
        n1, n2, n3, nnode = struct.unpack("<4i", h)
        nstate = 1
    else:
        raise RuntimeError(f"Unexpected RHO header record length: {len(h)} bytes")
    A = _lattice_from_al(rec.next())
    nr = n1 * n2 * n3
    nrn = nr // nnode
    data = np.zeros((nstate, nr))
    for s in range(nstate):
        for inode in range(nnode):
            r = rec.next()
            if 8 * nrn > len(r):
                raise RuntimeError("grid data: nbytes exceeds record length")
            data[s, inode * nrn:(inode + 1) * nrn] = np.frombuffer(r, dtype="<f8", count=nrn)
    return (n1, n2, n3), A, data


# --------------------------------------------------------------------------
#  OUT.EIGEN
# --------------------------------------------------------------------------
def read_eigen(path):
    """Return dict(islda, nkpt, nband, eig[islda, nkpt, nband] in Hartree,
    kvec[nkpt, 3], weight[nkpt])."""
    rec = _Records(path)
    h = rec.next()
    if len(h) == 28:
        islda, nkpt, nband, _nref, natom, nnode, is_so = struct.unpack("<7i", h)
    elif len(h) == 24:
        islda, nkpt, nband, _nref, natom, nnode = struct.unpack("<6i", h)
        is_so = 0
    else:
        raise RuntimeError(f"Unexpected EIGEN header record length: {len(h)} bytes")
    eig = np.zeros((islda, nkpt, nband))
    kvec = np.zeros((nkpt, 3))
    wt = np.zeros(nkpt)
    for s in range(islda):
        for k in range(nkpt):
            r = rec.next()
            if len(r) != 40:
                raise RuntimeError("kpt metadata: record size mismatch")
            isl, ik, w, ax, ay, az = struct.unpack("<2i4d", r)
            if isl != s + 1 or ik != k + 1:
                raise RuntimeError("EIGEN record order mismatch")
            wt[k] = w
            kvec[k] = (ax, ay, az)
            e = rec.next()
            if len(e) != 8 * nband:
                raise RuntimeError("eigenvalues: record size mismatch")
            eig[s, k] = np.frombuffer(e, dtype="<f8") / HARTREE_TO_EV
    return {"islda": islda, "nkpt": nkpt, "nband": nband, "natom": natom, "is_so": is_so,
            "eig": eig, "kvec": kvec, "weight": wt}


# --------------------------------------------------------------------------
#  atom.config
# --------------------------------------------------------------------------
def read_atom_config(path):
    """Return (A_bohr, Z[natom], frac[natom, 3]) with atoms grouped by
    ascending Z, file order kept within a species."""
    lines = [l.strip() for l in open(path).read().splitlines()]
    natom = int(lines[0].split()[0])
    i = next(j for j, l in enumerate(lines) if l.upper().startswith("LATTICE"))
    A = np.array([[float(x) for x in lines[i + 1 + r].split()[:3]] for r in range(3)]) / BOHR_RADIUS_ANGSTROM
    p = next(j for j, l in enumerate(lines) if l.upper().startswith("POSITION"))
    Z, pos = [], []
    for r in range(natom):
        f = lines[p + 1 + r].split()
        Z.append(int(f[0]))
        pos.append([float(x) for x in f[1:4]])
    Z = np.array(Z)
    order = np.argsort(Z, kind="stable")
    return A, Z[order], np.array(pos)[order]


# --------------------------------------------------------------------------
#  UPF v2 (norm-conserving) -> species tables
# --------------------------------------------------------------------------
def _floats(node):
    return np.array([float(x) for x in node.text.split()])


def read_upf(path):
    root = ET.parse(path).getroot()
    hdr = root.find("PP_HEADER").attrib
    mesh = root.find("PP_MESH")
    r = _floats(mesh.find("PP_R"))
    rab = _floats(mesh.find("PP_RAB"))
    nl = root.find("PP_NONLOCAL")
    nproj = int(hdr["number_of_proj"])
    betas, lll = [], []
    for i in range(1, nproj + 1):
        b = nl.find(f"PP_BETA.{i}")
        v = _floats(b)
        kb = int(b.attrib.get("cutoff_radius_index", len(v)))
        betas.append(v[:kb])                           # truncated (ncpp.cppm:215-219)
        lll.append(int(b.attrib["angular_momentum"]))
    dij = _floats(nl.find("PP_DIJ")).reshape(nproj, nproj)
    return {"element": hdr.get("element", "").strip(), "pseudo_type": hdr.get("pseudo_type", ""),
            "z_valence": float(hdr.get("z_valence", 0)), "r": r, "rab": rab, "betas": betas,
            "lll": np.array(lll, dtype=np.int32), "dion_ry": dij}


def simpson_weights(n, rab, uniform):
    """Weights w with sum(w * f) = composite Simpson's rule over the mesh
    index for n points (Simpson 3/8 on the first four when n is even),
    including the mesh derivative rab."""
    w = np.zeros(n)
    if n == 0:
        return w
    if n == 1:
        w[0] = 1.0
    elif n % 2 == 1:
        w[:] = 2.0
        w[1:-1:2] = 4.0
        w[0] = w[-1] = 1.0
        w /= 3.0
    elif n == 2:
        w[:] = 0.5
    else:
        w[:4] = np.array([1.0, 3.0, 3.0, 1.0]) * 3.0 / 8.0
        t = np.full(n - 3, 2.0)
        t[1:-1:2] = 4.0
        t[0] = t[-1] = 1.0
        w[3:] += t / 3.0
    return w * (rab[0] if uniform else rab[:n])


def simpson(f, rab, uniform):
    return float(np.dot(simpson_weights(len(f), rab, uniform), f))


def species_from_upf(upf, z, dq=synth.DQ, qmax=synth.QMAX) -> synth.Species:
    """sortByL + diagonalizeNonlocal + BetaqTables / DBetaqTables."""
    lll = upf["lll"]
    order = np.argsort(lll, kind="stable")
    lll = lll[order]
    betas = [upf["betas"][i] for i in order]
    D = upf["dion_ry"][np.ix_(order, order)] / RY_PER_HA
    nb = len(betas)
    maxcut = max(1, max(len(b) for b in betas))
    B = np.zeros((nb, maxcut))
    for i, b in enumerate(betas):
        B[i, :len(b)] = b
    lam = np.zeros(nb)
    rot = np.zeros_like(B)
    for l in sorted(set(lll.tolist())):
        idx = np.nonzero(lll == l)[0]
        w, U = np.linalg.eigh(D[np.ix_(idx, idx)])
        for k, ik in enumerate(idx):
            lam[ik] = w[k]
            rot[ik] = U[:, k] @ B[idx]
    r = upf["r"][:maxcut]
    rab = upf["rab"][:maxcut]
    uniform = bool(np.allclose(np.diff(upf["r"][:min(len(upf["r"]), 64)]), upf["r"][1] - upf["r"][0],
                               rtol=1e-9, atol=1e-12))
    nq = int(math.ceil(qmax / dq)) + 1 + 3
    qs = np.arange(nq) * dq
    tab = np.zeros((nb, nq))
    dtab = np.zeros((nb, nq))
    for ib in range(nb):
        l = int(lll[ib])
        f = rot[ib]
        n = len(f)
        while n > 0 and f[n - 1] == 0.0:
            n -= 1
        if n == 0:
            continue
        w = simpson_weights(n, rab, uniform)
        x = qs[:, None] * r[None, :n]
        jl = special.spherical_jn(l, x)
        jp = special.spherical_jn(l, x, derivative=True)
        tab[ib] = 4.0 * np.pi * (jl @ (w * f[:n] * r[:n]))
        dtab[ib] = 4.0 * np.pi * (jp @ (w * f[:n] * r[:n] ** 2))
    return synth.Species(z=z, lmax=int(lll.max()) if nb else -1, beta_l=lll.astype(np.int32),
                         lam=lam, beta_tab=tab, dbeta_tab=dtab, dq=dq, qmax=qmax)


def load_cell(atom_config, vr_path, upf_by_z: dict, n=None):
    """A synth.Cell from atom.config + VR + one UPF per element."""
    A, Z, frac = read_atom_config(atom_config)
    nvr, Avr, data = read_vr(vr_path)
    if n is not None and tuple(n) != tuple(nvr):
        raise RuntimeError("VR grid does not match the GKK grid")
    zs = sorted(set(Z.tolist()))
    species = np.array([zs.index(z) for z in Z], dtype=np.int32)
    sp = [species_from_upf(read_upf(upf_by_z[z]), z) for z in zs]
    return synth.Cell(A=A, n=tuple(nvr), vloc=np.ascontiguousarray(data[0]), tau=frac,
                      species=species, species_list=sp)


# Element symbols by atomic number (1-86), for matching atom.config's atomic
# numbers to the element names in the UPF headers.
ELEMENTS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br "
            "Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho "
            "Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn").split()


class PWmatHamiltonian:
    """The reference's file-based Hamiltonian interface
    (src/H_psi/hamiltonian.cppm:115-160, hamiltonian.cpp:131-153): loadATOM /
    loadNCPP / loadGKK / loadVR / loadEIGEN / loadWG with paths relative to
    base_dir, finalize(), loadFromDirectory(), the has*() queries, then
    at_k / gradient() / hessian() on the device Hamiltonian.  GKK k-point
    ikpt is made resident on the device the first time an operator asks for
    it.  NCPPs are matched to atom.config's atomic numbers by the element
    name of their UPF header, as the reference's ncpp(atomic_number) does
    (hamiltonian.cpp:199-206); finalize() raises if an atom has none."""

    def __init__(self, base_dir=".", device: int = 0, hbm_budget: int = 0):
        import os
        self._os = os
        self.base_dir = base_dir
        self.device = device
        self.hbm_budget = hbm_budget
        self.atom = self.gkk = self.eigen = self.wg = None
        self.vr_path = None
        self.ncpps = {}                 # element name -> parsed UPF
        self.H = None

    def _p(self, path):
        return path if self._os.path.isabs(path) else self._os.path.join(self.base_dir, path)

    # -- step-by-step loading --------------------------------------------------
    def loadATOM(self, path):
        self.atom = read_atom_config(self._p(path))

    def loadNCPP(self, path):
        upf = read_upf(self._p(path))
        self.ncpps[upf["element"]] = upf

    def loadGKK(self, path):
        self.gkk = GKKFile(self._p(path))

    def loadVR(self, path):
        read_vr(self._p(path))          # parsed now, so a bad file fails here
        self.vr_path = self._p(path)

    def loadEIGEN(self, path):
        self.eigen = read_eigen(self._p(path))

    def loadWG(self, path):
        self.wg = WGFile(self._p(path))

    def hasGKK(self):
        return self.gkk is not None

    def hasVR(self):
        return self.vr_path is not None

    def hasATOM(self):
        return self.atom is not None

    def hasEIGEN(self):
        return self.eigen is not None

    def hasWG(self):
        return self.wg is not None

    def finalize(self):
        """Species tables (sortByL, diagonalizeNonlocal, BetaqTables) and
        the device-resident cell; after every load, before any operator."""
        if self.atom is None or self.gkk is None or self.vr_path is None:
            raise RuntimeError("Hamiltonian: finalize() needs ATOM, GKK and VR")
        _, Z, _ = self.atom
        by_z = {}
        for z in sorted(set(int(v) for v in Z)):
            name = ELEMENTS[z - 1] if 1 <= z <= len(ELEMENTS) else str(z)
            if name not in self.ncpps:
                raise RuntimeError(f"Hamiltonian: no NCPP for element {name} (Z = {z})")
            by_z[z] = self.ncpps[name]
        A, Zs, frac = self.atom
        nvr, _, data = read_vr(self.vr_path)
        if tuple(nvr) != tuple(self.gkk.meta.n):
            raise RuntimeError("VR grid does not match the GKK grid")
        zs = sorted(by_z)
        species = np.array([zs.index(int(z)) for z in Zs], dtype=np.int32)
        cell = synth.Cell(A=A, n=tuple(nvr), vloc=np.ascontiguousarray(data[0]), tau=frac, species=species,
                          species_list=[species_from_upf(by_z[z], z) for z in zs])
        from .hamiltonian import Hamiltonian
        if self.H is not None:
            self.H.close()
        self.H = Hamiltonian(cell, device=self.device, hbm_budget=self.hbm_budget)
        return self

    def loadFromDirectory(self):
        """atom.config, the UPF files (UPF/ or base_dir), OUT.GKK, OUT.VR,
        and OUT.EIGEN / OUT.WG when present; then finalize()."""
        os = self._os
        self.loadATOM("atom.config")
        upf_dir = self._p("UPF") if os.path.isdir(self._p("UPF")) else self.base_dir
        for f in sorted(os.listdir(upf_dir)):
            if f.endswith((".UPF", ".upf")):
                self.loadNCPP(os.path.join(upf_dir, f))
        self.loadGKK("OUT.GKK")
        self.loadVR("OUT.VR")
        for name, load in (("OUT.EIGEN", self.loadEIGEN), ("OUT.WG", self.loadWG)):
            if os.path.exists(self._p(name)):
                load(name)
        return self.finalize()

    # -- operators -------------------------------------------------------------
    def _resident(self, ikpt):
        if self.H is None:
            raise RuntimeError("Hamiltonian: call finalize() before at_k() / gradient() / hessian()")
        if ikpt not in self.H._npw:
            self.H.load_kpoint(ikpt, self.gkk.kpoint(ikpt))
        return self.H

    def at_k(self, ikpt: int):
        return self._resident(ikpt).at_k(ikpt)

    def gradient(self):
        outer = self

        class _G:
            def at_k(self, ikpt):
                return outer._resident(ikpt).gradient().at_k(ikpt)
        return _G()

    def hessian(self, delta: float = 1e-3):
        outer = self

        class _Hs:
            def at_k(self, ikpt):
                return outer._resident(ikpt).hessian(delta).at_k(ikpt)
        return _Hs()

    def close(self):
        if self.H is not None:
            self.H.close()
            self.H = None
