"""PWmat / UPF readers (effmass_amd/pwmat.py) against the golden values the
reference's own tests hold: test/io/test_gkk.cpp:50-75, test/io/test_rho.cpp:28-36,
the k inference of test/test_wg_fft.cpp:64-80, and the Fourier-Bessel
closed form of test/math/test_fourier_bessel.cpp:225-228."""
import os

import numpy as np
import pytest

from effmass_amd import pwmat

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pwmat")


def test_gkk_golden_values():
    g = pwmat.GKKFile(os.path.join(G, "aln_OUT.GKK"))
    assert g.meta.n == (20, 20, 32) and g.meta.nnode == 1 and g.nkpt == 10
    kp = g.kpoint(0)
    assert kp.npw == 1737
    assert abs(kp.kin[0] - 0.7529635525206351) < 1e-12
    assert abs(kp.kin[1] - 2.258890654552743) < 1e-12
    assert abs(kp.kin[2] - 5.270744861626122) < 1e-12
    assert abs(kp.K[0, 0] - -1.062753653146566) < 1e-12
    assert abs(kp.K[1, 0] - -2.125507306293132) < 1e-12
    assert abs(kp.K[0, 1] - -0.6135811093611817) < 1e-12
    assert abs(kp.K[0, 2]) < 1e-15
    # kinetic = |K|^2 / 2 (GKK.cppm:652-671) and the integer view reproduces K
    assert np.abs(kp.kin - 0.5 * (kp.K ** 2).sum(1)).max() < 1e-10 * (1 + kp.kin.max())
    B = 2 * np.pi * np.linalg.inv(g.meta.A).T
    assert np.abs((kp.gidx + kp.kfrac) @ B - kp.K).max() < 1e-10
    with pytest.raises(IndexError):
        g.kpoint(10)
    with pytest.raises(IndexError):
        g.kpoint(-1)


def test_gkk_kpoint_inference_vs_eigen():
    for gkk, eig in (("si8_OUT2.GKK", "si8_OUT2.EIGEN"), ("si8_OUT1.GKK", "si8_OUT1.EIGEN"),
                     ("aln_OUT.GKK", "aln_OUT.EIGEN")):
        g = pwmat.GKKFile(os.path.join(G, gkk))
        e = pwmat.read_eigen(os.path.join(G, eig))
        assert e["nkpt"] == g.nkpt
        B = 2 * np.pi * np.linalg.inv(g.meta.A).T
        for ik in range(g.nkpt):
            kp = g.kpoint(ik)
            assert np.abs(kp.kfrac @ B - e["kvec"][ik]).max() < 1e-10, (gkk, ik)
    assert pwmat.GKKFile(os.path.join(G, "si8_OUT2.GKK")).kpoint(0).npw == 6451


def _raw_records(path):
    import struct
    buf = open(path, "rb").read()
    o, out = 0, []
    while o < len(buf):
        (n,) = struct.unpack_from("<i", buf, o)
        out.append(buf[o + 4:o + 4 + n])
        o += 8 + n
    return out


def test_gkk_two_node_merge():
    """OUT1.GKK holds each k-point on 2 nodes (test/data_nonscf, the file
    the reference's gradient test loads).  loadKPoint concatenates the
    nodes' first ngtotnod entries in node order and negates the stored
    -(k+G) (src/io/GKK.cppm:525-549): npw is the sum over nodes, and the
    merged G-set is the whole sphere |k+G|^2 <= Ecut, each G once."""
    path = os.path.join(G, "si8_OUT1.GKK")
    g = pwmat.GKKFile(path)
    assert g.meta.nnode == 2 and g.nkpt == 7 and g.meta.n == (36, 36, 36)
    assert g.ngtotnod[0].tolist() == [3245, 3222]
    recs = _raw_records(path)
    B = 2 * np.pi * np.linalg.inv(g.meta.A).T
    r = 4                                            # header, Ecut, AL, ngtotnod
    for ik in range(g.nkpt):
        kp = g.kpoint(ik)
        ng = g.ngtotnod[ik]
        assert kp.npw == int(ng.sum())
        parts = []
        for inode in range(2):
            m = int(ng[inode])
            kin, x, y, z = (np.frombuffer(recs[r + a], "<f8", count=m) for a in range(4))
            parts.append((kin, -np.stack([x, y, z], 1)))
            r += 4
        assert np.array_equal(kp.kin, np.concatenate([p[0] for p in parts]))
        assert np.array_equal(kp.K, np.concatenate([p[1] for p in parts]))
        # the merged set is the full sphere, each integer G exactly once
        assert len({tuple(v) for v in kp.gidx.tolist()}) == kp.npw
        ax = [np.arange(-18, 18)] * 3
        gg = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
        k2 = (((gg + kp.kfrac) @ B) ** 2).sum(1)
        # the header's Ecut bounds the kinetic energy |K|^2 / 2 (Hartree)
        inside = int((0.5 * k2 <= g.meta.ecut * (1 + 1e-9)).sum())
        assert inside == kp.npw, (ik, inside, kp.npw)
    assert r == len(recs)


def test_vr_golden_values():
    n, A, d = pwmat.read_vr(os.path.join(G, "aln_OUT.VR"))
    assert n == (20, 20, 32) and d.shape == (1, 12800)
    assert abs(d[0, 0] - 0.0932307015284446) < 1e-9
    assert abs(d[0, 1] - 0.1119047176083688) < 1e-9


def test_eigen_and_atom_config():
    e = pwmat.read_eigen(os.path.join(G, "aln_OUT.EIGEN"))
    assert e["nkpt"] == 10 and e["nband"] == 26 and e["eig"].shape == (1, 10, 26)
    A, Z, frac = pwmat.read_atom_config(os.path.join(G, "aln_atom.config"))
    assert list(Z) == [7, 7, 13, 13]                 # ascending Z, stable (ATOM.cppm:290-316)
    A8, Z8, f8 = pwmat.read_atom_config(os.path.join(G, "si8_atom.config"))
    assert np.allclose(A8, np.eye(3) * 5.43102018 / 0.52917721067)


def test_upf_d_projectors():
    """Ge-spd-high.PD04 (test/data_upf): norm-conserving with l = 0, 1, 2
    projector pairs; sortByL keeps them, the l = 2 tables vanish like q^2."""
    u = pwmat.read_upf(os.path.join(G, "Ge-spd-high.PD04.PBE.UPF"))
    assert u["pseudo_type"] == "NC" and list(u["lll"]) == [0, 0, 1, 1, 2, 2]
    sp = pwmat.species_from_upf(u, 32)
    assert sp.lmax == 2 and list(sp.beta_l) == [0, 0, 1, 1, 2, 2]
    D = u["dion_ry"] / 2
    for l in (0, 1, 2):
        idx = np.nonzero(u["lll"] == l)[0]
        assert np.allclose(np.sort(sp.lam[sp.beta_l == l]), np.linalg.eigvalsh(D[np.ix_(idx, idx)]))
    d = sp.beta_tab[sp.beta_l == 2]
    q = np.arange(sp.nq) * sp.dq
    assert np.all(d[:, 0] == 0.0)
    ratio = d[:, 1:6] / q[1:6] ** 2                    # beta_2(q) ~ q^2 near 0
    assert np.abs(ratio / ratio[:, :1] - 1).max() < 1e-3


def test_upf_and_tables():
    u = pwmat.read_upf(os.path.join(G, "Si.SG15.PBE.UPF"))
    assert list(u["lll"]) == [0, 0, 1, 1] and u["dion_ry"].shape == (4, 4)
    sp = pwmat.species_from_upf(u, 14)
    assert sp.nq == 2004 and sp.beta_tab.shape == (4, 2004)
    # D is used through its l-blocks only: lambda are their eigenvalues (Ha)
    D = u["dion_ry"] / 2
    for l, idx in ((0, [0, 1]), (1, [2, 3])):
        assert np.allclose(np.sort(sp.lam[sp.beta_l == l]), np.linalg.eigvalsh(D[np.ix_(idx, idx)]))
    # tables: l > 0 vanish at q = 0; dbeta/dq is the derivative of beta(q)
    assert np.all(sp.beta_tab[sp.beta_l > 0, 0] == 0.0)
    fd = (sp.beta_tab[:, 2:] - sp.beta_tab[:, :-2]) / (2 * sp.dq)
    assert np.abs(fd - sp.dbeta_tab[:, 1:-1]).max() < 1e-4 * np.abs(sp.dbeta_tab).max()


def test_fourier_bessel_closed_form():
    # UPF stores r*beta(r); with beta(r) = r e^{-r^2} and l = 1 the table is
    # 4 pi int r^3 e^{-r^2} j_1(q r) dr = pi^{3/2} q e^{-q^2/4} / 2, whose
    # q-derivative at 0 is pi^{3/2} / 2
    r = np.arange(0, 2001) * 0.005
    rab = np.full_like(r, 0.005)
    upf = {"lll": np.array([1], np.int32), "betas": [r * r * np.exp(-r * r)], "dion_ry": np.eye(1) * 2,
           "r": r, "rab": rab}
    sp = pwmat.species_from_upf(upf, 1)
    q = np.arange(sp.nq) * sp.dq
    exact = np.pi ** 1.5 * q * np.exp(-q * q / 4) / 2
    assert np.abs(sp.beta_tab[0] - exact).max() < 1e-8
    assert abs(sp.dbeta_tab[0, 0] - np.pi ** 1.5 / 2) < 1e-8


def _rec(payload: bytes) -> bytes:
    import struct
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def test_wg_hand_built_records(tmp_path):
    """A two-node, two-band OUT.WG assembled record by record in the layout
    of src/io/WG.cppm:251-277 (header), 280-307 (ngtotnod), 320-337 (per k,
    node-major then band) and 398-406 (first ngtotnod coefficients valid)."""
    import struct
    # AL(3,3) column-major: AL(:, i) is lattice vector i, so the flat record
    # is vector 0, vector 1, vector 2 (Angstrom); a hexagonal cell pins the order
    al = np.array([[3.11, 0.0, 0.0], [-1.555, 2.6933, 0.0], [0.0, 0.0, 4.98]])
    ng = [3, 2]                                                       # per node, mg_nx = 4
    rng = np.random.default_rng(0)
    c = rng.standard_normal((2, 5)) + 1j * rng.standard_normal((2, 5))
    body = _rec(struct.pack("<9i", 16, 18, 20, 2, 4, 2, 1, 0, 1))
    body += _rec(struct.pack("<d", 30.0))
    body += _rec(al.astype("<f8").tobytes())
    body += _rec(struct.pack("<3i", 2, *ng))
    for node, (a, b) in enumerate([(0, 3), (3, 5)]):
        for band in range(2):
            rec = np.zeros(4, "<c16")
            rec[:b - a] = c[band, a:b]
            rec[b - a:] = 99.0                                         # padding is ignored
            body += _rec(rec.tobytes())
    p = tmp_path / "OUT.WG"
    p.write_bytes(body)
    w = pwmat.WGFile(str(p))
    m = w.meta
    assert m.n == (16, 18, 20) and m.nband == 2 and m.mg_nx == 4 and m.nnode == 2 and m.nkpt == 1
    assert m.is_so == 0 and m.ecut == 30.0 and list(m.ng_per_kpt) == [5]
    assert np.allclose(m.A, al / pwmat.BOHR_RADIUS_ANGSTROM)
    up, down = w.band(0, 1)
    assert down is None and np.array_equal(up, c[1])
    assert np.array_equal(w.bands(0), c)
    with pytest.raises(IndexError):
        w.band(1, 0)
    with pytest.raises(IndexError):
        w.band(0, 2)
    # a broken trailing marker is a record-length mismatch
    bad = bytearray(body)
    bad[-1] ^= 0x01
    p.write_bytes(bytes(bad))
    with pytest.raises(RuntimeError):
        pwmat.WGFile(str(p))


def test_wg_spin_orbit_halves_and_round_trip(tmp_path):
    """Spin-orbit records hold the up half then the down half, each mg_nx
    long; the header counts both halves (src/io/WG.cppm:229-249, 266-268,
    297-299).  Also a multi-k, multi-node round trip through write_wg."""
    rng = np.random.default_rng(1)
    A = np.diag([10.26, 10.26, 10.26])
    ks = []
    splits = [[4, 3], [5, 1], [2, 6]]
    for s in splits:
        ng = sum(s)
        ks.append((rng.standard_normal((3, ng)) + 1j * rng.standard_normal((3, ng)),
                   rng.standard_normal((3, ng)) + 1j * rng.standard_normal((3, ng))))
    p = str(tmp_path / "OUT.WG.so")
    pwmat.write_wg(p, (24, 24, 24), A, 40.0, ks, node_split=splits, mg_nx=7)
    w = pwmat.WGFile(p)
    assert w.meta.is_so == 1 and w.meta.mg_nx == 7 and w.meta.nkpt == 3
    assert list(w.meta.ng_per_kpt) == [7, 6, 8]
    assert np.allclose(w.meta.A, A)
    import struct
    raw = open(p, "rb").read()
    assert struct.unpack_from("<9i", raw, 4)[4] == 14             # file header: both halves
    for ik in range(3):
        for ib in range(3):
            up, dn = w.band(ik, ib)
            assert np.array_equal(up, ks[ik][0][ib]) and np.array_equal(dn, ks[ik][1][ib])
        assert np.array_equal(w.bands(ik, [2, 0], spin="down"), ks[ik][1][[2, 0]])
    # without spin-orbit
    q = str(tmp_path / "OUT.WG")
    plain = [k[0] for k in ks]
    pwmat.write_wg(q, (24, 24, 24), A, 40.0, plain, node_split=splits)
    w2 = pwmat.WGFile(q)
    assert w2.meta.is_so == 0
    for ik in range(3):
        assert np.array_equal(w2.bands(ik), plain[ik])
    with pytest.raises(ValueError):
        w2.bands(0, spin="down")


def _standard_dir(tmp_path, upf_subdir=True):
    """The Si8 fixtures under the names loadFromDirectory reads."""
    import os
    d = tmp_path / "si8"
    d.mkdir()
    for src, dst in (("si8_atom.config", "atom.config"), ("si8_OUT2.GKK", "OUT.GKK"), ("si8_IN.VR", "OUT.VR"),
                     ("si8_OUT2.EIGEN", "OUT.EIGEN")):
        os.symlink(os.path.join(G, src), d / dst)
    u = d / "UPF" if upf_subdir else d
    u.mkdir(exist_ok=True)
    os.symlink(os.path.join(G, "Si.SG15.PBE.UPF"), u / "Si.SG15.PBE.UPF")
    return d


def test_file_loader_steps_and_queries(tmp_path):
    """The reference's step-by-step loading names (hamiltonian.cppm:121-160)
    read the same data as the readers; has*() report what was loaded."""
    d = _standard_dir(tmp_path)
    h = pwmat.PWmatHamiltonian(str(d))
    assert not (h.hasATOM() or h.hasGKK() or h.hasVR() or h.hasEIGEN() or h.hasWG())
    h.loadATOM("atom.config")
    h.loadNCPP("UPF/Si.SG15.PBE.UPF")
    h.loadGKK("OUT.GKK")
    h.loadVR("OUT.VR")
    h.loadEIGEN("OUT.EIGEN")
    assert h.hasATOM() and h.hasGKK() and h.hasVR() and h.hasEIGEN() and not h.hasWG()
    assert list(h.ncpps) == ["Si"]
    assert h.gkk.meta.n == (36, 36, 36)
    A, Z, _ = h.atom
    assert pwmat.ELEMENTS[int(Z[0]) - 1] == "Si"


def test_file_loader_finalize_needs_every_element(tmp_path):
    """finalize() refuses a cell with an atom whose element has no NCPP
    (the reference's NCPPAtomCoverage check), before any device work."""
    d = _standard_dir(tmp_path)
    h = pwmat.PWmatHamiltonian(str(d))
    h.loadATOM("atom.config")
    h.loadGKK("OUT.GKK")
    h.loadVR("OUT.VR")
    with pytest.raises(RuntimeError, match="no NCPP for element Si"):
        h.finalize()
    h2 = pwmat.PWmatHamiltonian(str(d))
    with pytest.raises(RuntimeError, match="needs ATOM, GKK and VR"):
        h2.finalize()
