"""Batched ragged applies (effm_hpsi_multi / effm_dhdk_multi) and hipGraph
replay (EFFM_F_GRAPH), through the C-ABI.

Tolerances: the FP64 batch equals the per-k FP64 apply to 1e-13 relative
per band (the split-K partial sums may be grouped differently) and the
oracle to 1e-12; a graph replay equals the eager apply bit for bit (same
kernels, same arguments)."""
import sys

import numpy as np
import pytest

import oracle
from conftest import workload
from effmass_amd import synth
from effmass_amd.hamiltonian import GradHPsi, Hamiltonian

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def band_rel(got, ref):
    num = np.linalg.norm(got - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    return float((num / den).max())


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_multi(nk, grid=(36, 36, 36), nb=6, seed=11):
    """fcc Si2 at nk random k-points: ragged npw, quick for the oracle."""
    rng = np.random.default_rng(seed)
    ks = [rng.uniform(-0.5, 0.5, 3) for _ in range(nk)]
    return synth.custom(*synth.fcc_primitive(10.26), grid, 30.0, ks, nb, seed=seed)


def _load(wl):
    H = Hamiltonian(wl.cell)
    for ik, kp in enumerate(wl.kpts):
        H.load_kpoint(ik, kp)
    return H


def test_multi_kgrid_shapes_equals_single_k():
    """64-atom cell, 5 k-points of the 4x4x2 grid (ragged npw), 8 bands: the
    batched launch against one effm_hpsi per k-point and the oracle."""
    torch = _torch()
    wl = workload("si64_kgrid", nb=8, nk=5)
    assert len({kp.npw for kp in wl.kpts}) > 1
    log(f"si64 k-grid shapes npw {[kp.npw for kp in wl.kpts]}")
    H = _load(wl)
    psis = [torch.from_numpy(p).cuda() for p in wl.psi]
    outs = [torch.zeros_like(p) for p in psis]
    H.apply_multi(list(range(5)), psis, outs, overwrite=True)
    for ik in range(5):
        single = torch.zeros_like(psis[ik])
        H.at_k(ik)(psis[ik], single, overwrite=True)
        torch.cuda.synchronize()
        e = band_rel(outs[ik].cpu().numpy(), single.cpu().numpy())
        log(f"k {ik}: batch vs single {e:.2e}")
        assert e <= 1e-13
    ref = oracle.hpsi(wl.cell, wl.kpts[3], wl.psi[3], nthreads=8)
    assert band_rel(outs[3].cpu().numpy(), ref) <= 1e-12
    # mixed: batch vs single within FP32 rounding
    p32 = [p.to(torch.complex64) for p in psis]
    o32 = [torch.zeros_like(p) for p in p32]
    H.apply_multi(list(range(5)), p32, o32, prec="mixed", overwrite=True)
    s32 = torch.zeros_like(p32[1])
    H.at_k(1)(p32[1], s32, prec="mixed", overwrite=True)
    torch.cuda.synchronize()
    assert band_rel(o32[1].cpu().numpy().astype(np.complex128), s32.cpu().numpy().astype(np.complex128)) <= 1e-5
    H.close()


def test_multi_over_sixteen_kpoints_and_accumulate():
    """18 k-points (two launches of at most 16), accumulating onto a start
    value, every k-point against the oracle."""
    torch = _torch()
    wl = _small_multi(18)
    H = _load(wl)
    psis = [torch.from_numpy(p).cuda() for p in wl.psi]
    base = [torch.from_numpy(np.full(p.shape, 0.25 - 0.5j)).cuda() for p in wl.psi]
    outs = [b.clone() for b in base]
    H.apply_multi(list(range(18)), psis, outs)
    torch.cuda.synchronize()
    worst = 0.0
    for ik in range(18):
        ref = oracle.hpsi(wl.cell, wl.kpts[ik], wl.psi[ik]) + base[ik].cpu().numpy()
        worst = max(worst, band_rel(outs[ik].cpu().numpy(), ref))
    log(f"18 k-points worst rel err {worst:.2e}")
    assert worst <= 1e-12
    H.close()


def test_multi_generic_grid_kernels():
    """A grid without compile-time plans (28 x 30 x 32) takes the generic
    batched kernels (one grid slice per segment)."""
    torch = _torch()
    ks = [np.array([0.13, 0.0, -0.27]), np.array([0.5, 0.25, 0.0]), np.zeros(3)]
    wl = synth.custom(*synth.fcc_primitive(10.26), (28, 30, 32), 12.0, ks, 4, seed=5)
    H = _load(wl)
    psis = [torch.from_numpy(p).cuda() for p in wl.psi]
    outs = [torch.zeros_like(p) for p in psis]
    H.apply_multi([0, 1, 2], psis, outs, overwrite=True)
    torch.cuda.synchronize()
    for ik in range(3):
        assert band_rel(outs[ik].cpu().numpy(), oracle.hpsi(wl.cell, wl.kpts[ik], wl.psi[ik])) <= 1e-12
    H.close()


def test_dhdk_multi_equals_single_k():
    torch = _torch()
    wl = _small_multi(4)
    H = _load(wl)
    psis = [torch.from_numpy(p).cuda() for p in wl.psi]
    ox = [torch.zeros_like(p) for p in psis]
    oz = [torch.zeros_like(p) for p in psis]
    H.gradient_multi([0, 1, 2, 3], psis, GradHPsi(ox, None, oz), overwrite=True)
    for ik in range(4):
        sx, sz = torch.zeros_like(psis[ik]), torch.zeros_like(psis[ik])
        H.gradient().at_k(ik)(psis[ik], GradHPsi(sx, None, sz), overwrite=True)
        torch.cuda.synchronize()
        assert band_rel(ox[ik].cpu().numpy(), sx.cpu().numpy()) <= 1e-13
        assert band_rel(oz[ik].cpu().numpy(), sz.cpu().numpy()) <= 1e-13
    gx, gy, gz = oracle.dhdk(wl.cell, wl.kpts[2], wl.psi[2])
    assert band_rel(ox[2].cpu().numpy(), gx) <= 1e-12
    assert band_rel(oz[2].cpu().numpy(), gz) <= 1e-12
    H.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fused_hpsi_dhdk_equals_separate_calls(prec):
    """effm_hpsi_dhdk_multi (one projection GEMM for H|psi> and dH/dk|psi>)
    against effm_hpsi_multi + effm_dhdk_multi: overwrite, then accumulate,
    then a graph replay; FP64 also against the oracle.  Tolerance: FP64
    1e-13 relative per band (the split-K sums of the fused projection are
    grouped like the gradient's), mixed 1e-6."""
    torch = _torch()
    wl = _small_multi(5, nb=8, seed=23)
    H = _load(wl)
    cdt = torch.complex128 if prec == "fp64" else torch.complex64
    tol = 1e-13 if prec == "fp64" else 1e-6
    ks = list(range(5))
    psis = [torch.from_numpy(p).cuda().to(cdt) for p in wl.psi]
    hs = [torch.zeros_like(p) for p in psis]
    g = GradHPsi([torch.zeros_like(p) for p in psis], None, [torch.zeros_like(p) for p in psis])
    rh = [torch.zeros_like(p) for p in psis]
    rg = GradHPsi([torch.zeros_like(p) for p in psis], None, [torch.zeros_like(p) for p in psis])

    def check(tag):
        torch.cuda.synchronize()
        for ik in ks:
            for got, ref in ((hs[ik], rh[ik]), (g.x[ik], rg.x[ik]), (g.z[ik], rg.z[ik])):
                err = band_rel(got.cpu().numpy().astype(np.complex128), ref.cpu().numpy().astype(np.complex128))
                assert err <= tol, (tag, ik, err)
        log(f"{prec} {tag}: fused == separate")

    H.apply_with_gradient_multi(ks, psis, hs, g, prec=prec, overwrite=True)
    H.apply_multi(ks, psis, rh, prec=prec, overwrite=True)
    H.gradient_multi(ks, psis, rg, prec=prec, overwrite=True)
    check("overwrite")
    H.apply_with_gradient_multi(ks, psis, hs, g, prec=prec)
    H.apply_multi(ks, psis, rh, prec=prec)
    H.gradient_multi(ks, psis, rg, prec=prec)
    check("accumulate")
    for _ in range(3):
        H.apply_with_gradient_multi(ks, psis, hs, g, prec=prec, overwrite=True, graph=True)
    H.apply_multi(ks, psis, rh, prec=prec, overwrite=True)
    H.gradient_multi(ks, psis, rg, prec=prec, overwrite=True)
    check("graph replay")
    if prec == "fp64":
        hp = oracle.hpsi(wl.cell, wl.kpts[3], wl.psi[3])
        gx, gy, gz = oracle.dhdk(wl.cell, wl.kpts[3], wl.psi[3])
        assert band_rel(hs[3].cpu().numpy(), hp) <= 1e-12
        assert band_rel(g.x[3].cpu().numpy(), gx) <= 1e-12
        assert band_rel(g.z[3].cpu().numpy(), gz) <= 1e-12
    H.close()


def test_graph_replay_matches_eager():
    """Captured apply sequences replayed several times, with eager work and
    new psi contents in between; every replay equals the eager result."""
    torch = _torch()
    wl = _small_multi(3, nb=5, seed=21)
    H = _load(wl)
    psis = [torch.from_numpy(p).cuda() for p in wl.psi]
    outs = [torch.zeros_like(p) for p in psis]
    eager = [torch.zeros_like(p) for p in psis]
    g = [torch.zeros_like(p) for p in psis]
    rng = np.random.default_rng(3)
    for rep in range(4):
        for p in psis:               # same buffers, new contents
            p.copy_(torch.from_numpy(rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)))
        H.apply_multi([0, 1, 2], psis, g, overwrite=True, graph=True)
        H.apply_multi([0, 1, 2], psis, eager, overwrite=True)
        torch.cuda.synchronize()
        for a, b in zip(g, eager):
            assert torch.equal(a, b), f"replay {rep} differs"
        # single-k H|psi> and dH/dk replays
        h1 = torch.zeros_like(psis[1])
        H.at_k(1)(psis[1], h1, overwrite=True, graph=True)
        h2 = torch.zeros_like(psis[1])
        H.at_k(1)(psis[1], h2, overwrite=True)
        torch.cuda.synchronize()
        assert torch.equal(h1, h2)
        dg = [torch.zeros_like(p) for p in psis]
        de = [torch.zeros_like(p) for p in psis]
        H.gradient_multi([0, 1, 2], psis, GradHPsi(None, dg, None), overwrite=True, graph=True)
        H.gradient_multi([0, 1, 2], psis, GradHPsi(None, de, None), overwrite=True)
        torch.cuda.synchronize()
        for a, b in zip(dg, de):
            assert torch.equal(a, b), f"dH/dk replay {rep} differs"
    # accumulate mode through a graph: out += H psi on every replay
    acc = torch.zeros_like(psis[0])
    for rep in range(3):
        H.at_k(0)(psis[0], acc, graph=True)
    one = torch.zeros_like(psis[0])
    H.at_k(0)(psis[0], one, overwrite=True)
    torch.cuda.synchronize()
    assert band_rel(acc.cpu().numpy(), 3 * one.cpu().numpy()) <= 1e-14
    # a new allocation (another k-point) drops the cached graphs; the next
    # graph call captures again and is still right
    H.load_kpoint(3, wl.kpts[0])
    H.apply_multi([0, 1, 2], psis, g, overwrite=True, graph=True)
    torch.cuda.synchronize()
    for ik in range(3):
        assert band_rel(g[ik].cpu().numpy(), oracle.hpsi(wl.cell, wl.kpts[ik], psis[ik].cpu().numpy())) <= 1e-12
    H.close()


def test_autotune_choices_keep_parity():
    """The autotuner picks among the candidate shapes; every later apply,
    whatever was picked, still matches the oracle."""
    torch = _torch()
    wl = workload("si64_cubic", nb=24)
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, wl.kpts[0])
    for prec in ("fp64", "mixed"):
        pb, gw = H.autotune(0, 24, prec)
        sh = H.shapes(prec)
        log(f"{prec}: {sh}")
        assert pb in (4, 8, 16) and gw in (512, 768, 1024, 1536, 2048)
        assert (sh["pencil_bands"], sh["gemm_wgs"]) == (pb, gw)
        assert sh["pencil_z"] in (16, 18) and sh["apply_tile"] == 64     # 24 bands: no tile-width choice
    psi = torch.from_numpy(wl.psi[0][:24]).cuda()
    h = torch.zeros_like(psi)
    H.at_k(0)(psi, h, overwrite=True)
    torch.cuda.synchronize()
    ref = oracle.hpsi(wl.cell, wl.kpts[0], wl.psi[0][:4], nthreads=8)
    assert band_rel(h[:4].cpu().numpy(), ref) <= 1e-12
    H.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_kgrid32_hpsi_dhdk_graph(prec):
    """The 32-k-point batch at its own shape (64-atom cell, 4x4x2 grid,
    ragged npw 36.8k-37.1k, two launch groups of 16 k-points; 4 nch = 2048
    projection rows over several tiles per segment): effm_hpsi_dhdk_multi
    replayed as a hipGraph equals, at every k-point, the single-k effm_hpsi
    + effm_dhdk (FP64 1e-13; mixed 1e-5), and k-points 0 and 31 equal the
    oracle (FP64 1e-12, mixed 2e-5).  A second replay after new psi contents
    equals an eager call bit for bit.  hamiltonian_gradient.cpp:116-294."""
    torch = _torch()
    nb = 16
    wl = workload("si64_kgrid", nb=nb)
    nk = len(wl.kpts)
    assert nk == 32 and len({kp.npw for kp in wl.kpts}) > 8
    cdt = torch.complex128 if prec == "fp64" else torch.complex64
    H = _load(wl)
    ld = max(kp.npw for kp in wl.kpts)
    # one ld for the whole batch (rows padded past npw), as a caller's pool
    psis = []
    for p in wl.psi:
        t = torch.zeros((nb, ld), dtype=cdt, device="cuda")
        t[:, :p.shape[1]] = torch.from_numpy(p).to(cdt)
        psis.append(t)
    hs = [torch.zeros_like(p) for p in psis]
    gs = GradHPsi(*[[torch.zeros_like(p) for p in psis] for _ in range(3)])
    ks = list(range(nk))
    H.apply_with_gradient_multi(ks, psis, hs, gs, prec=prec, overwrite=True, graph=True)
    torch.cuda.synchronize()
    tol_single = 1e-13 if prec == "fp64" else 1e-5
    worst = 0.0
    for ik in ks:
        npw = wl.kpts[ik].npw
        sh = torch.zeros_like(psis[ik])
        H.at_k(ik)(psis[ik], sh, prec=prec, overwrite=True)
        sg = [torch.zeros_like(psis[ik]) for _ in range(3)]
        H.gradient().at_k(ik)(psis[ik], GradHPsi(*sg), prec=prec, overwrite=True)
        torch.cuda.synchronize()
        pairs = [(hs[ik], sh)] + [(getattr(gs, d)[ik], sg[a]) for a, d in enumerate("xyz")]
        for got, ref in pairs:
            worst = max(worst, band_rel(got[:, :npw].cpu().numpy().astype(np.complex128),
                                        ref[:, :npw].cpu().numpy().astype(np.complex128)))
    log(f"{prec}: 32 k-points batch+graph vs single-k worst {worst:.2e}")
    assert worst <= tol_single
    tol = 1e-12 if prec == "fp64" else 2e-5
    for ik in (0, 31):
        kp, psi = wl.kpts[ik], wl.psi[ik][:8]
        ref = oracle.hpsi(wl.cell, kp, psi, nthreads=8)
        gref = oracle.dhdk(wl.cell, kp, psi)
        e = band_rel(hs[ik][:8, :kp.npw].cpu().numpy().astype(np.complex128), ref)
        eg = max(band_rel(getattr(gs, d)[ik][:8, :kp.npw].cpu().numpy().astype(np.complex128), gref[a])
                 for a, d in enumerate("xyz"))
        log(f"{prec}: k {ik} vs oracle: H {e:.2e}, dH/dk {eg:.2e}")
        assert e <= tol and eg <= tol
    # new psi contents, same buffers: the replay equals an eager call
    for p in psis:
        p.mul_(complex(0.6, -0.8))
    H.apply_with_gradient_multi(ks, psis, hs, gs, prec=prec, overwrite=True, graph=True)
    eh = [torch.zeros_like(p) for p in psis]
    eg_ = GradHPsi(*[[torch.zeros_like(p) for p in psis] for _ in range(3)])
    H.apply_with_gradient_multi(ks, psis, eh, eg_, prec=prec, overwrite=True, graph=False)
    torch.cuda.synchronize()
    for ik in (0, 15, 16, 31):
        assert torch.equal(hs[ik], eh[ik])
        for d in "xyz":
            assert torch.equal(getattr(gs, d)[ik], getattr(eg_, d)[ik])
    H.close()


def test_multi_split_plane_grid_and_graph():
    """A 112^3 grid (FP64 planes beyond the LDS: split-plane kernels) in a
    batch of 3 ragged k-points: equals the oracle at 1e-12, and a graph
    replay after new psi contents equals the eager batch bit for bit."""
    torch = _torch()
    A, tau = synth.diamond_cubic(10.26)
    ks = [np.array([0.1, -0.05, 0.2]), np.array([0.5, 0.0, 0.25]), np.zeros(3)]
    wl = synth.custom(A, tau, (112, 112, 112), 20.0, ks, 3, seed=112)
    H = _load(wl)
    psis = [torch.from_numpy(p).cuda() for p in wl.psi]
    outs = [torch.zeros_like(p) for p in psis]
    H.apply_multi([0, 1, 2], psis, outs, overwrite=True, graph=True)
    torch.cuda.synchronize()
    for ik in range(3):
        ref = oracle.hpsi(wl.cell, wl.kpts[ik], wl.psi[ik], nthreads=8)
        e = band_rel(outs[ik].cpu().numpy(), ref)
        log(f"112^3 k {ik}: {e:.2e}")
        assert e <= 1e-12
    for p in psis:
        p.mul_(complex(0.0, 1.0))
    H.apply_multi([0, 1, 2], psis, outs, overwrite=True, graph=True)
    eager = [torch.zeros_like(p) for p in psis]
    H.apply_multi([0, 1, 2], psis, eager, overwrite=True)
    torch.cuda.synchronize()
    for ik in range(3):
        assert torch.equal(outs[ik], eager[ik])
    H.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_every_launch_shape_gives_the_same_result(prec):
    """Every candidate launch shape gives the default shape's results on the
    64-atom cell (90^3, 512 projector channels) with 40 bands (two 32-wide
    apply tiles, the second one partial): the x-pencil band group 4 / 8 / 16
    and z width 16 / 18 bit for bit (every band and z column is transformed
    on its own), the apply tile width 64 / 32 bit for bit (each output keeps
    its k order) for H|psi> and dH/dk, the split-K targets to 1e-13 (FP64) /
    1e-6 (mixed) relative per band (partial sums grouped differently); the
    default shape equals the oracle (FP64 1e-12, mixed 2e-5).  Values outside
    the candidate lists are refused and change nothing."""
    torch = _torch()
    nb = 40
    wl = workload("si64_cubic", nb=nb)
    kp = wl.kpts[0]
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    p = torch.from_numpy(wl.psi[0][:nb]).to(dt).cuda()

    def run():
        h = torch.zeros_like(p)
        H.at_k(0)(p, h, prec=prec)
        g = [torch.zeros_like(p) for _ in range(3)]
        H.gradient().at_k(0)(p, GradHPsi(*g), prec=prec)
        torch.cuda.synchronize()
        return h.cpu().numpy(), [x.cpu().numpy() for x in g]

    default = H.shapes(prec)
    assert default == {"pencil_bands": 8, "pencil_z": 18 if prec == "fp64" else 16, "gemm_wgs": 1024,
                       "apply_tile": 64}
    h0, g0 = run()
    ref = oracle.hpsi(wl.cell, kp, wl.psi[0][:4], nthreads=8)
    tol = 1e-12 if prec == "fp64" else 2e-5
    assert band_rel(h0[:4], ref) <= tol
    for pb in (4, 8, 16):
        for pz in (16, 18):
            H.set_shapes(prec, pencil_bands=pb, pencil_z=pz)
            h, _ = run()
            log(f"{prec} pencil {pb} bands x {pz} z: max |diff| {np.abs(h - h0).max():.1e}")
            assert np.array_equal(h, h0)
    H.set_shapes(prec, **default)
    H.set_shapes(prec, apply_tile=32)
    assert H.shapes(prec)["apply_tile"] == 32
    h, g = run()
    log(f"{prec} apply tile 32: max |diff| H {np.abs(h - h0).max():.1e}, "
        f"dH/dk {max(np.abs(g[a] - g0[a]).max() for a in range(3)):.1e}")
    assert np.array_equal(h, h0) and all(np.array_equal(g[a], g0[a]) for a in range(3))
    H.set_shapes(prec, apply_tile=64)
    stol = 1e-13 if prec == "fp64" else 1e-6
    for w in (512, 768, 1536, 2048):
        H.set_shapes(prec, gemm_wgs=w)
        h, g = run()
        e = max([band_rel(h, h0)] + [band_rel(g[a], g0[a]) for a in range(3)])
        log(f"{prec} split-K target {w}: {e:.1e}")
        assert e <= stol
    H.set_shapes(prec, **default)
    from effmass_amd._lib import EffmInvalidArgument
    for bad in ({"pencil_bands": 32}, {"pencil_z": 20}, {"gemm_wgs": 1000}, {"apply_tile": 16}):
        with pytest.raises(EffmInvalidArgument):
            H.set_shapes(prec, **bad)
        assert H.shapes(prec) == default
    H.close()


@pytest.mark.parametrize("grid", [(28, 30, 32), (120, 120, 120)])
def test_autotune_on_generic_and_split_grids(grid):
    """The autotuner on cells outside the compile-time plans: the runtime-plan
    kernels (28 x 30 x 32) and the split-plane kernels (120^3, whose FP64
    plane exceeds the LDS), with 40 bands so that the apply tile width is
    tuned as well.  After tuning, H|psi> equals the default shapes' result
    (FP64 1e-13, mixed 1e-6 relative per band: the split-K partial sums may
    be grouped differently) and the oracle on two bands (FP64 1e-12, mixed
    2e-5)."""
    torch = _torch()
    A, tau = synth.diamond_cubic(10.26)
    nb = 40
    wl = synth.custom(A, tau, grid, 25.0, [np.array([0.1, -0.05, 0.2])], nb, seed=7 + sum(grid))
    kp = wl.kpts[0]
    assert kp.npw >= nb
    ref = oracle.hpsi(wl.cell, kp, wl.psi[0][:2], nthreads=8)
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    for prec, dt, tol, stol in (("fp64", torch.complex128, 1e-12, 1e-13), ("mixed", torch.complex64, 2e-5, 1e-6)):
        p = torch.from_numpy(wl.psi[0]).to(dt).cuda()
        h0 = torch.zeros_like(p)
        H.at_k(0)(p, h0, prec=prec, overwrite=True)
        H.autotune(0, nb, prec)
        sh = H.shapes(prec)
        h = torch.zeros_like(p)
        H.at_k(0)(p, h, prec=prec, overwrite=True)
        torch.cuda.synchronize()
        got, base = h.cpu().numpy(), h0.cpu().numpy()
        e_shape = band_rel(got, base)
        e = band_rel(got[:2].astype(np.complex128), ref)
        log(f"grid {grid} {prec} npw {kp.npw}: tuned {sh}; vs default shapes {e_shape:.1e}, vs oracle {e:.2e}")
        assert e_shape <= stol and e <= tol
    H.close()
