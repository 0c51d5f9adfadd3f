"""Parity of the gfx950 H|psi> / dH/dk path (through the C-ABI) with the
FP64 oracle.  Tolerances: FP64 device path <= 1e-12 relative per band;
mixed path energies within 1e-6 Ry (5e-7 Ha) of the FP64 oracle."""
import copy
import sys

import numpy as np
import pytest

import oracle
from conftest import workload
from effmass_amd import synth, _lib
from effmass_amd.hamiltonian import GradHPsi, Hamiltonian

pytestmark = pytest.mark.gpu

RY_TOL_HA = 0.5e-6


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def band_rel(got, ref):
    num = np.linalg.norm(got - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    return float((num / den).max())


def run_h(wl, ik=0, nb=None, prec="fp64", dev=True):
    torch = _torch()
    H = Hamiltonian(wl.cell)
    H.load_kpoint(ik, wl.kpts[ik])
    psi = wl.psi[ik][:nb] if nb else wl.psi[ik]
    if prec == "mixed":
        psi = psi.astype(np.complex64)
    if dev:
        p = torch.from_numpy(psi).cuda()
        h = torch.zeros_like(p)
        H.at_k(ik)(p, h, prec=prec)
        torch.cuda.synchronize()
        out = h.cpu().numpy()
    else:
        out = np.zeros_like(psi)
        H.at_k(ik)(psi, out, prec=prec)
    H.close()
    return out


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_si2_fcc_fp64():
    wl = workload("si2_fcc")
    got = run_h(wl)
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0])
    e = band_rel(got, ref)
    log(f"si2 fcc fp64 rel err {e:.2e}")
    assert e <= 1e-12


def test_two_species_offgamma_fp64():
    wl = synth.custom(*synth.fcc_primitive(10.26), (36, 36, 36), 30.0,
                      [np.array([0.2, -0.1, 0.35])], 5, seed=11,
                      species_idx=[0, 1],
                      species_list=[synth.make_species(synth.SplitMix64(1)),
                                    synth.make_species(synth.SplitMix64(2), z=7, ls=(0, 1, 1),
                                                       widths=(0.3, 0.6))])
    got = run_h(wl)
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0])
    assert band_rel(got, ref) <= 1e-12


def test_si8_cubic_fp64_subset():
    wl = workload("si8_cubic", nb=24)
    got = run_h(wl)
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0], nthreads=8)
    e = band_rel(got, ref)
    log(f"si8 cubic fp64 rel err {e:.2e}")
    assert e <= 1e-12


def test_si64_cubic_fp64_subset():
    wl = workload("si64_cubic", nb=16)
    got = run_h(wl)
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0], nthreads=8)
    e = band_rel(got, ref)
    log(f"si64 cubic fp64 rel err {e:.2e}")
    assert e <= 1e-12


def test_si64_cubic_mixed_energies():
    wl = workload("si64_cubic", nb=16)
    got = run_h(wl, prec="mixed")
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0], nthreads=8)
    psi = wl.psi[0]
    e_ref = np.real(np.sum(psi.conj() * ref, axis=1)) / np.sum(np.abs(psi) ** 2, axis=1)
    p32 = psi.astype(np.complex64).astype(np.complex128)
    e_mix = np.real(np.sum(p32.conj() * got, axis=1)) / np.sum(np.abs(p32) ** 2, axis=1)
    de = np.abs(e_mix - e_ref).max()
    log(f"si64 cubic mixed max |dE| {de:.2e} Ha, rel hpsi err {band_rel(got, ref):.2e}")
    assert de <= RY_TOL_HA
    assert band_rel(got, ref) <= 2e-5


def test_accumulate_overwrite_and_host_path():
    torch = _torch()
    wl = workload("si2_fcc")
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, wl.kpts[0])
    psi = torch.from_numpy(wl.psi[0]).cuda()
    base = torch.from_numpy(np.arange(wl.psi[0].size).reshape(wl.psi[0].shape) * (1 - 1j)).cuda()
    h = base.clone()
    H.at_k(0)(psi, h)                                   # accumulate (reference semantics)
    h2 = base.clone()
    H.at_k(0)(psi, h2, overwrite=True)
    torch.cuda.synchronize()
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0])
    assert np.abs(h.cpu().numpy() - base.cpu().numpy() - ref).max() <= 1e-14 * np.abs(base.cpu().numpy()).max()
    assert band_rel(h2.cpu().numpy(), ref) <= 1e-12
    hh = np.zeros_like(wl.psi[0])
    H.at_k(0)(wl.psi[0], hh)                            # host pointers
    assert band_rel(hh, ref) <= 1e-12
    # one band at a time equals the batch (Callable(psi, hpsi) on a span)
    hb = torch.zeros_like(psi[0])
    H.at_k(0)(psi[3], hb)
    torch.cuda.synchronize()
    assert band_rel(hb.cpu().numpy()[None], ref[3:4]) <= 1e-12
    H.close()


def test_dhdk_fp64():
    torch = _torch()
    wl = synth.custom(*synth.fcc_primitive(10.26), (36, 36, 36), 30.0,
                      [np.array([0.1, 0.05, -0.2])], 6, seed=5)
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, wl.kpts[0])
    psi = torch.from_numpy(wl.psi[0]).cuda()
    outs = [torch.zeros_like(psi) for _ in range(3)]
    H.gradient().at_k(0)(psi, GradHPsi(*outs))
    torch.cuda.synchronize()
    ref = oracle.dhdk(wl.cell, wl.kpts[0], wl.psi[0])
    for a in range(3):
        e = band_rel(outs[a].cpu().numpy(), ref[a])
        log(f"dhdk dir {a} rel err {e:.2e}")
        assert e <= 1e-12
    # a skipped middle direction (empty span, hamiltonian.cppm:53-65) does
    # not shift the others: x and z still equal the oracle's
    x = torch.zeros_like(psi)
    z = torch.zeros_like(psi)
    H.gradient().at_k(0)(psi, GradHPsi(x=x, y=None, z=z))
    torch.cuda.synchronize()
    assert band_rel(x.cpu().numpy(), ref[0]) <= 1e-12
    assert band_rel(z.cpu().numpy(), ref[2]) <= 1e-12
    H.close()


def test_dhdk_mixed():
    torch = _torch()
    wl = workload("si2_fcc")
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, wl.kpts[0])
    psi = torch.from_numpy(wl.psi[0].astype(np.complex64)).cuda()
    outs = [torch.zeros_like(psi) for _ in range(3)]
    H.gradient().at_k(0)(psi, GradHPsi(*outs), prec="mixed")
    torch.cuda.synchronize()
    ref = oracle.dhdk(wl.cell, wl.kpts[0], wl.psi[0])
    for a in range(3):
        assert band_rel(outs[a].cpu().numpy(), ref[a]) <= 2e-5
    H.close()


def test_rayleigh_fp64_and_mixed():
    torch = _torch()
    wl = workload("si2_fcc")
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, wl.kpts[0])
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0])
    psi = wl.psi[0]
    e_ref = np.real(np.sum(psi.conj() * ref, 1)) / np.sum(np.abs(psi) ** 2, 1)
    E = H.rayleigh(torch.from_numpy(psi).cuda(), torch.from_numpy(ref).cuda()).cpu().numpy()
    assert np.abs(E - e_ref).max() <= 1e-13 * np.abs(e_ref).max()
    E32 = H.rayleigh(psi.astype(np.complex64), ref.astype(np.complex64), prec="mixed")
    assert np.abs(E32 - e_ref).max() <= 1e-6
    H.close()


def test_error_paths():
    torch = _torch()
    wl = workload("si2_fcc")
    H = Hamiltonian(wl.cell)
    with pytest.raises(_lib.EffmOutOfRange):
        H.at_k(0)                                           # no k-point loaded
    kp = wl.kpts[0]
    bad = copy.copy(kp)
    bad.kin = kp.kin.copy()
    bad.kin[5] += 1e-6
    with pytest.raises(_lib.EffmRuntimeError):              # GKK kinetic mismatch
        H.load_kpoint(0, bad)
    far = copy.copy(kp)
    far.K = kp.K * 4.0                                      # q beyond the 20 bohr^-1 table
    far.kin = 0.5 * (far.K ** 2).sum(1)
    with pytest.raises(_lib.EffmDomainError):
        H.load_kpoint(0, far)
    H.load_kpoint(0, kp)
    resident, _ = H.mem_info()                              # cell + k-point data
    H.close()
    proj = wl.cell.nchannels() * kp.npw * 16                # FP64 projector matrix
    plane_bytes = 36 * 36 * 16
    p = torch.from_numpy(wl.psi[0]).cuda()
    # the FFT workspace is chunked to fit the budget: room for about two
    # bands' planes still gives the right answer ...
    nplanes = len(np.unique(kp.gidx[:, 0] % 36))
    Hs = Hamiltonian(wl.cell, hbm_budget=resident + proj + 2 * nplanes * plane_bytes + (1 << 20))
    try:
        Hs.load_kpoint(0, kp)
        h = torch.zeros_like(p)
        Hs.at_k(0)(p, h)
        torch.cuda.synchronize()
        assert band_rel(h.cpu().numpy(), oracle.hpsi(wl.cell, kp, wl.psi[0])) <= 1e-12
    finally:
        Hs.close()
    # an unknown precision is refused, not read as complex64
    Hr = Hamiltonian(wl.cell)
    try:
        lib = _lib.load()
        E = torch.zeros(2, dtype=torch.float64, device="cuda")
        rc = lib.effm_rayleigh(Hr.ctx, kp.npw, 2, p.data_ptr(), kp.npw, p.data_ptr(), kp.npw, 7, E.data_ptr(), 0)
        assert rc == _lib.EFFM_ERR_INVALID
    finally:
        Hr.close()
    # ... and no room for one band's planes fails loudly
    with pytest.raises(_lib.EffmOutOfMemory):
        Hs = Hamiltonian(wl.cell, hbm_budget=resident + proj + nplanes * plane_bytes // 4)
        try:
            Hs.load_kpoint(0, kp)
            Hs.at_k(0)(p, torch.zeros_like(p))
        finally:
            Hs.close()


def test_si64_cubic_full_hermitian_and_linear():
    """All 256 bands at the 64-atom size: size-independent properties."""
    torch = _torch()
    wl = workload("si64_cubic")
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, wl.kpts[0])
    psi = torch.from_numpy(wl.psi[0]).cuda()
    h = torch.zeros_like(psi)
    H.at_k(0)(psi, h)
    M = (psi.conj() @ h.T).cpu().numpy()                     # <psi_m|H|psi_n>
    herm = np.abs(M - M.conj().T).max() / np.abs(M).max()
    log(f"si64 cubic hermiticity {herm:.2e}")
    assert herm <= 1e-12
    a, b = 0.3 - 0.7j, -1.1 + 0.2j
    comb = (a * psi[:8] + b * psi[8:16]).contiguous()
    hc = torch.zeros_like(comb)
    H.at_k(0)(comb, hc)
    lin = (hc - (a * h[:8] + b * h[8:16])).abs().max().item() / h[:16].abs().max().item()
    assert lin <= 1e-12
    # mixed path energies over all 256 bands within 1e-6 Ry of the FP64 path
    p32 = psi.to(torch.complex64)
    h32 = torch.zeros_like(p32)
    H.at_k(0)(p32, h32, prec="mixed")
    E64 = H.rayleigh(psi, h).cpu().numpy()
    E32 = H.rayleigh(p32, h32, prec="mixed").cpu().numpy()
    de = np.abs(E32 - E64).max()
    log(f"si64 cubic all-band mixed |dE| max {de:.2e} Ha")
    assert de <= RY_TOL_HA
    H.close()


def test_multi_k_ragged():
    torch = _torch()
    wl = workload("si64_kgrid", nb=4, nk=3)
    H = Hamiltonian(wl.cell)
    for ik, kp in enumerate(wl.kpts):
        H.load_kpoint(ik, kp)
    psis = [torch.from_numpy(p).cuda() for p in wl.psi]
    outs = [torch.zeros_like(p) for p in psis]
    H.apply_multi([0, 1, 2], psis, outs)
    torch.cuda.synchronize()
    assert len({kp.npw for kp in wl.kpts}) > 1                # ragged npw
    for ik in range(3):
        ref = oracle.hpsi(wl.cell, wl.kpts[ik], wl.psi[ik], nthreads=8)
        assert band_rel(outs[ik].cpu().numpy(), ref) <= 1e-12
    H.close()


@pytest.mark.parametrize("grid", [(28, 30, 32), (42, 45, 40), (24, 20, 36), (48, 49, 50),
                                  (15, 15, 15), (16, 16, 16), (18, 18, 18), (20, 20, 20), (25, 25, 25),
                                  (40, 24, 24), (45, 45, 45), (48, 25, 25), (48, 45, 45), (16, 40, 40),
                                  (22, 26, 33), (26, 22, 39),
                                  (27, 32, 32), (32, 27, 27), (96, 80, 80), (80, 96, 96)])
def test_fft_radix_plans(grid):
    """Grids whose axes use radices 2..12 (incl. 7) in the runtime LDS plan,
    and every compile-time two-stage plan (square planes and pencils of
    15, 16, 18, 20, 24, 25, 27, 32, 40, 45, 48, 80, 96; 36 and 90 in the
    tests above): prime-factor order where the two
    radices are coprime (15, 20, 40, 45), Cooley-Tukey otherwise, and the
    two mixed between the pencil and the plane axes ((40, 24, 24): prime-
    factor pencil, Cooley-Tukey planes; (48, 45, 45), (16, 40, 40): the
    reverse), and axes with the prime radices 11 and 13 (22, 26, 33, 39).
    FP64 against the oracle at 1e-12, the mixed path within 1e-5."""
    wl = synth.custom(*synth.fcc_primitive(10.26), grid, 12.0,
                      [np.array([0.13, 0.0, -0.27])], 4, seed=sum(grid))
    got = run_h(wl)
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0])
    e = band_rel(got, ref)
    em = band_rel(run_h(wl, prec="mixed"), ref)           # the FP32 kernels of the same plans
    log(f"grid {grid} rel err {e:.2e}, mixed {em:.2e}")
    assert e <= 1e-12
    assert em <= 1e-5


@pytest.mark.parametrize("grid,ecut,nb", [((24, 24, 24), 2.0, 3), ((24, 24, 24), 12.0, 67),
                                          ((36, 36, 36), 30.0, 300)])
def test_odd_sizes(grid, ecut, nb):
    """A sphere of a few plane waves, band counts that are not multiples of
    the GEMM tiles (64) or of the pencil band groups, and more bands than
    four column tiles: FP64 H|psi> and dH/dk against the oracle at 1e-12;
    the mixed path's Rayleigh quotients within 1e-6 Ry of the FP64 path's
    and its dH/dk within 2e-5."""
    torch = _torch()
    wl = synth.custom(*synth.fcc_primitive(10.26), grid, ecut, [np.array([0.1, -0.2, 0.05])], nb,
                      seed=int(ecut * 10) + nb)
    kp = wl.kpts[0]
    log(f"grid {grid} ecut {ecut} npw {kp.npw} nb {nb}")
    got = run_h(wl)
    ref = oracle.hpsi(wl.cell, kp, wl.psi[0])
    assert band_rel(got, ref) <= 1e-12
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    psi = torch.from_numpy(wl.psi[0]).cuda()
    h = torch.zeros_like(psi)
    H.at_k(0)(psi, h)
    p32 = psi.to(torch.complex64)
    h32 = torch.zeros_like(p32)
    H.at_k(0)(p32, h32, prec="mixed")
    de = np.abs(H.rayleigh(p32, h32, prec="mixed").cpu().numpy() - H.rayleigh(psi, h).cpu().numpy()).max()
    log(f"  mixed |dE| max {de:.2e} Ha")
    assert de <= RY_TOL_HA
    # dH/dk (one projection launch over the projector and k-gradient rows):
    # FP64 at 1e-12, mixed at 2e-5 against the oracle
    ref = oracle.dhdk(wl.cell, kp, wl.psi[0])
    outs = [torch.zeros_like(psi) for _ in range(3)]
    H.gradient().at_k(0)(psi, GradHPsi(*outs))
    outs32 = [torch.zeros_like(p32) for _ in range(3)]
    H.gradient().at_k(0)(p32, GradHPsi(*outs32), prec="mixed")
    torch.cuda.synchronize()
    for a in range(3):
        assert band_rel(outs[a].cpu().numpy(), ref[a]) <= 1e-12
        assert band_rel(outs32[a].cpu().numpy(), ref[a]) <= 2e-5
    H.close()


def _raw_dion_wl(ls, seed, widths=(0.25, 0.5), nb=6):
    """2-atom fcc cell, off Gamma, whose species has a random symmetric dion
    over all channels (non-diagonal l-blocks, non-zero cross-l entries)."""
    sp = synth.make_species(synth.SplitMix64(seed), ls=ls, widths=widths)
    return synth.custom(*synth.fcc_primitive(10.26), (36, 36, 36), 30.0, [np.array([0.15, -0.05, 0.2])], nb,
                        seed=seed + 1, species_list=[sp])


def _h_and_grad(wl, prec):
    torch = _torch()
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, wl.kpts[0])
    psi = wl.psi[0].astype(np.complex64 if prec == "mixed" else np.complex128)
    p = torch.from_numpy(psi).cuda()
    h = torch.zeros_like(p)
    H.at_k(0)(p, h, prec=prec)
    g = [torch.zeros_like(p) for _ in range(3)]
    H.gradient().at_k(0)(p, GradHPsi(*g), prec=prec)
    torch.cuda.synchronize()
    out = h.cpu().numpy(), [x.cpu().numpy() for x in g]
    H.close()
    return out


@pytest.mark.parametrize("ls", [(0, 1, 2, 3), (0, 0, 1, 1, 2, 2)])
def test_high_l_projectors(ls):
    """Projectors up to l = 3 on the device: the three-term Legendre advance,
    its dtheta form and the m >= 2 gradient seeds of the device Y_lm
    (k_nonlocal.hip ylm_eval) against the oracle; FP64 H|psi> and dH/dk at
    1e-12, mixed at 2e-5."""
    wl = _raw_dion_wl(ls, seed=91 + len(ls), widths=(0.3, 0.45))
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0])
    gref = oracle.dhdk(wl.cell, wl.kpts[0], wl.psi[0])
    for prec, tol in (("fp64", 1e-12), ("mixed", 2e-5)):
        h, g = _h_and_grad(wl, prec)
        e = band_rel(h, ref)
        eg = max(band_rel(g[a], gref[a]) for a in range(3))
        log(f"ls {ls} {prec}: H {e:.2e}, dH/dk {eg:.2e}")
        assert e <= tol and eg <= tol


def test_raw_dion_dense_on_device():
    """The device H with a non-diagonal dion (cross-l entries included)
    equals the dense H whose nonlocal part is sum_ij |p_i> D_ij <p_j| over
    the UNROTATED closed-form projectors with cross-l D dropped
    (src/pseudo/ncpp_advance.cppm:65-153), at the table-interpolation
    tolerance."""
    import refmath
    wl = _raw_dion_wl((0, 0, 1, 1, 2), seed=57)
    kp = wl.kpts[0]
    Hd = refmath.dense_h(wl.cell, kp, nonlocal_=False) + refmath.dense_vnl_raw(wl.cell, kp)
    ref = wl.psi[0] @ Hd.T
    h, _ = _h_and_grad(wl, "fp64")
    e = band_rel(h, ref)
    log(f"raw dion dense pin {e:.2e}")
    assert e <= 1e-8


def test_ge_d_projector_upf():
    """The reference's norm-conserving Ge-spd-high.PD04 UPF (l = 0, 1, 2
    projector pairs; test/data_upf) on a 2-atom Ge fcc cell: FP64 H|psi> and
    dH/dk equal the oracle at 1e-12, mixed at 2e-5."""
    import os
    from effmass_amd import pwmat
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pwmat")
    sp = pwmat.species_from_upf(pwmat.read_upf(os.path.join(g, "Ge-spd-high.PD04.PBE.UPF")), 32)
    wl = synth.custom(*synth.fcc_primitive(10.69), (36, 36, 36), 30.0, [np.array([0.25, 0.1, -0.15])], 6,
                      seed=32, species_list=[sp])
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0])
    gref = oracle.dhdk(wl.cell, wl.kpts[0], wl.psi[0])
    for prec, tol in (("fp64", 1e-12), ("mixed", 2e-5)):
        h, gr = _h_and_grad(wl, prec)
        e = band_rel(h, ref)
        eg = max(band_rel(gr[a], gref[a]) for a in range(3))
        log(f"Ge PD04 {prec}: H {e:.2e}, dH/dk {eg:.2e}")
        assert e <= tol and eg <= tol


@pytest.mark.parametrize("grid", [(108, 108, 108), (120, 120, 120), (128, 128, 128), (100, 112, 104)])
def test_large_grids(grid):
    """Grids whose FP64 n2 x n3 plane exceeds the 160 KiB LDS take the
    split-plane kernels (row and column passes through the workspace);
    their FP32 planes still fit and take the whole-plane runtime-plan
    kernels.  FP64 H|psi> against the oracle at 1e-12; mixed within 2e-5 and
    its Rayleigh quotients within 1e-6 Ry of the FP64 path's."""
    torch = _torch()
    A, tau = synth.diamond_cubic(10.26)
    wl = synth.custom(A, tau, grid, 25.0, [np.array([0.1, -0.05, 0.2])], 3, seed=sum(grid))
    kp = wl.kpts[0]
    ref = oracle.hpsi(wl.cell, kp, wl.psi[0], nthreads=8)
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    psi = torch.from_numpy(wl.psi[0]).cuda()
    h = torch.zeros_like(psi)
    H.at_k(0)(psi, h)
    p32 = psi.to(torch.complex64)
    h32 = torch.zeros_like(p32)
    H.at_k(0)(p32, h32, prec="mixed")
    torch.cuda.synchronize()
    e = band_rel(h.cpu().numpy(), ref)
    e32 = band_rel(h32.cpu().numpy().astype(np.complex128), ref)
    de = np.abs(H.rayleigh(p32, h32, prec="mixed").cpu().numpy() - H.rayleigh(psi, h).cpu().numpy()).max()
    log(f"grid {grid} npw {kp.npw}: FP64 {e:.2e}, mixed {e32:.2e}, |dE| {de:.2e} Ha")
    H.close()
    assert e <= 1e-12 and e32 <= 2e-5 and de <= RY_TOL_HA


@pytest.mark.parametrize("nb", [1, 2, 15, 16, 17, 24, 32, 33])
def test_narrow_band_counts(nb):
    """Band counts at and around the 16- and 32-wide GEMM tiles (one band
    per call, as the reference's callers apply H; the shrinking row blocks
    of a Chebyshev filter) on the 8-atom cell (64 projector channels, npw in
    the thousands: the narrow tiles and the many-split reduce both run):
    FP64 H|psi> and dH/dk at 1e-12 against the oracle, mixed at 2e-5; the
    same bands through a 40-band call (64-wide tiles) agree with the narrow
    call to 1e-13 (FP64)."""
    torch = _torch()
    wl = workload("si8_cubic", nb=40)
    kp = wl.kpts[0]
    psi = wl.psi[0][:nb]
    ref = oracle.hpsi(wl.cell, kp, psi, nthreads=8)
    gref = oracle.dhdk(wl.cell, kp, psi)
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    for prec, dt, tol in (("fp64", torch.complex128, 1e-12), ("mixed", torch.complex64, 2e-5)):
        p = torch.from_numpy(psi).to(dt).cuda()
        h = torch.zeros_like(p)
        H.at_k(0)(p, h, prec=prec)
        g = [torch.zeros_like(p) for _ in range(3)]
        H.gradient().at_k(0)(p, GradHPsi(*g), prec=prec)
        torch.cuda.synchronize()
        e = band_rel(h.cpu().numpy(), ref)
        eg = max(band_rel(g[a].cpu().numpy(), gref[a]) for a in range(3))
        log(f"nb {nb} {prec}: hpsi {e:.2e} dhdk {eg:.2e}")
        assert e <= tol and eg <= tol
    full = torch.from_numpy(wl.psi[0]).cuda()
    hf = torch.zeros_like(full)
    H.at_k(0)(full, hf)
    p = full[:nb].contiguous()
    h1 = torch.zeros_like(p)
    H.at_k(0)(p, h1)
    torch.cuda.synchronize()
    assert band_rel(h1.cpu().numpy(), hf[:nb].cpu().numpy()) <= 1e-13
    H.close()


def test_apply_gemm_quarter_tiles():
    """The apply GEMM's per-CU remainder tiles as 64x16 quarter tiles: 150
    bands on the 8-atom cell give 101 x 3 = 303 tiles of 64 x 64, i.e. 256
    whole ones and 47 run as quarter tiles, the last column tile partial
    (bands 128-149).  H|psi> and dH/dk from one 150-band call against one
    call per band (the separate 64 x 16 tile kernel) at 1e-13 (FP64) / 1e-6
    (mixed), and the partial column tile's bands against the oracle
    (FP64 1e-12, mixed 2e-5)."""
    torch = _torch()
    nb = 150
    wl = workload("si8_cubic", nb=nb)
    kp = wl.kpts[0]
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    tail = slice(128, nb)
    ref = oracle.hpsi(wl.cell, kp, wl.psi[0][tail], nthreads=8)
    gref = oracle.dhdk(wl.cell, kp, wl.psi[0][tail])
    for prec, dt, tol_call, tol_ref in (("fp64", torch.complex128, 1e-13, 1e-12),
                                        ("mixed", torch.complex64, 1e-6, 2e-5)):
        p = torch.from_numpy(wl.psi[0]).to(dt).cuda()
        h = torch.zeros_like(p)
        H.at_k(0)(p, h, prec=prec)
        g = [torch.zeros_like(p) for _ in range(3)]
        H.gradient().at_k(0)(p, GradHPsi(*g), prec=prec)
        h1 = torch.zeros_like(p)
        g1 = [torch.zeros_like(p) for _ in range(3)]
        for b in range(nb):
            H.at_k(0)(p[b:b + 1], h1[b:b + 1], prec=prec)
            H.gradient().at_k(0)(p[b:b + 1], GradHPsi(*(x[b:b + 1] for x in g1)), prec=prec)
        torch.cuda.synchronize()
        hb, h1n = h.cpu().numpy(), h1.cpu().numpy()
        e_call = max(band_rel(hb, h1n), max(band_rel(g[a].cpu().numpy(), g1[a].cpu().numpy()) for a in range(3)))
        e_ref = max(band_rel(hb[tail], ref), max(band_rel(g[a].cpu().numpy()[tail], gref[a]) for a in range(3)))
        log(f"quarter tiles {prec}: batch vs per-band {e_call:.2e}, tail vs oracle {e_ref:.2e}")
        assert e_call <= tol_call and e_ref <= tol_ref
    H.close()


def test_thousand_bands():
    """1000 bands of the 8-atom cell (npw 6451, 48^3) in one call: the
    widest launches in this suite (16 band tiles of 64 per GEMM tile row,
    125 x-pencil band groups of 8, 1000 x 23 plane items).  FP64 H|psi> and
    dH/dk against the oracle on the first, middle and last bands (1e-12),
    mixed at 2e-5; the 1000-band call equals two 500-band calls to 1e-13
    (FP64, relative per band: the split-K partial sums may be grouped
    differently) and 1e-6 (mixed)."""
    torch = _torch()
    nb = 1000
    wl = workload("si8_cubic", nb=nb)
    kp = wl.kpts[0]
    pick = [0, 499, 500, 999]
    ref = oracle.hpsi(wl.cell, kp, wl.psi[0][pick], nthreads=8)
    gref = oracle.dhdk(wl.cell, kp, wl.psi[0][[0, 999]])
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    for prec, dt, tol, stol in (("fp64", torch.complex128, 1e-12, 1e-13), ("mixed", torch.complex64, 2e-5, 1e-6)):
        p = torch.from_numpy(wl.psi[0]).to(dt).cuda()
        h = torch.zeros_like(p)
        H.at_k(0)(p, h, prec=prec)
        halves = [torch.zeros_like(p[:500]) for _ in range(2)]
        H.at_k(0)(p[:500].contiguous(), halves[0], prec=prec)
        H.at_k(0)(p[500:].contiguous(), halves[1], prec=prec)
        g = [torch.zeros_like(p) for _ in range(3)]
        H.gradient().at_k(0)(p, GradHPsi(*g), prec=prec)
        torch.cuda.synchronize()
        hn = h.cpu().numpy()
        e = band_rel(hn[pick].astype(np.complex128), ref)
        es = band_rel(np.concatenate([x.cpu().numpy() for x in halves]), hn)
        eg = max(band_rel(g[a].cpu().numpy()[[0, 999]].astype(np.complex128), gref[a]) for a in range(3))
        log(f"1000 bands {prec}: hpsi vs oracle {e:.2e}, dH/dk vs oracle {eg:.2e}, vs two 500-band calls {es:.1e}")
        assert e <= tol and eg <= tol and es <= stol
    H.close()
