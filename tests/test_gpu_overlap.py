"""effm_overlap (FP64-reduced split-K overlaps conj(A) B^T) and LOBPCG on
the device with it, through the C-ABI.  Tolerances: FP64 overlaps equal a
complex128 matmul to 1e-13 relative (max element); complex64 inputs to 1e-6;
device LOBPCG eigenvalues agree with dense diagonalisation of the closed-form H to 1e-7 Ha (table interpolation)."""
import numpy as np
import pytest

import refmath
from conftest import workload
from effmass_amd import eigensolver
from effmass_amd.hamiltonian import Hamiltonian

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 101, 777), (256, 242, 24001), (64, 64, 0), (130, 3, 5000),
                                   (512, 16, 30001), (512, 17, 30001), (200, 1, 40000)])
def test_overlap_matches_matmul(m, n, k):
    torch = _torch()
    wl = workload("si2_fcc")
    H = Hamiltonian(wl.cell)
    g = torch.Generator(device="cuda").manual_seed(m * 1000 + n + k)
    A = torch.randn(m, k, dtype=torch.complex128, device="cuda", generator=g)
    B = torch.randn(n, k, dtype=torch.complex128, device="cuda", generator=g)
    C = H.overlap(A, B)
    ref = A.conj() @ B.T
    scale = max(1.0, float(ref.abs().max()))
    assert float((C - ref).abs().max()) <= 1e-13 * scale * max(1.0, np.sqrt(k) / 10)
    C32 = H.overlap(A.to(torch.complex64), B.to(torch.complex64))
    assert float((C32.to(torch.complex128) - ref).abs().max()) <= 1e-6 * scale * max(1.0, np.sqrt(k) / 10)
    H.close()


def test_device_lobpcg_with_engine_overlaps():
    """LOBPCG on the device (H|psi> and overlaps through the engine) on the
    2-atom cell, against dense diagonalisation of H_GG'."""
    torch = _torch()
    wl = workload("si2_fcc")
    kp = wl.kpts[0]
    ref = np.linalg.eigvalsh(refmath.dense_h(wl.cell, kp))[:8]
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    kin = torch.from_numpy(kp.kin).cuda()
    X0 = eigensolver.random_start(12, kp.npw, kin, seed=4, device="cuda")
    op = H.at_k(0)
    res = eigensolver.lobpcg(lambda V, HV: op(V, HV, overwrite=True), X0, kin, nconv=8, tol=1e-9,
                             gram=H.overlap)
    assert res.converged == 8
    # dense H_GG' uses closed-form projectors: the table interpolation of the
    # engine differs at 1e-8 (see the oracle pins), so compare at 1e-7 Ha
    assert np.abs(res.E[:8].cpu().numpy() - ref).max() < 1e-7
    G = (res.X.conj() @ res.X.T).cpu().numpy()
    assert np.abs(G - np.eye(12)).max() < 1e-12
    H.close()


@pytest.mark.parametrize("n", [1, 255, 256 * 1024 + 7, 3 * 4096 * 1000])
def test_lincomb_matches_torch(n):
    """effm_lincomb z <- a z + b y + c x against the same complex128 torch
    expression: FP64 to 1e-15 relative; complex64 storage (FP64 arithmetic)
    to 1e-6."""
    torch = _torch()
    wl = workload("si2_fcc")
    H = Hamiltonian(wl.cell)
    g = torch.Generator(device="cuda").manual_seed(n)
    z, y, x = (torch.randn(n, dtype=torch.complex128, device="cuda", generator=g) for _ in range(3))
    a, b, c = 1.7, -0.31, 2.5e-3
    ref = a * z + b * y + c * x
    for dt, tol in ((torch.complex128, 1e-15), (torch.complex64, 1e-6)):
        zz = z.to(dt, copy=True)           # .to() of the same dtype would alias z
        H.lincomb(zz, a, y.to(dt).contiguous(), b, x.to(dt).contiguous(), c)
        torch.cuda.synchronize()
        assert float((zz.to(torch.complex128) - ref).abs().max()) <= tol * max(1.0, float(ref.abs().max()))
    with pytest.raises(TypeError):
        H.lincomb(z, 1.0, y.to(torch.complex64), 1.0, x, 1.0)
    H.close()


def test_chfsi_with_engine_lincomb():
    """ChFSI with the filter recurrence in effm_lincomb, on the 2-atom cell
    in both precisions, against dense diagonalisation of H_GG': FP64 to
    1e-7 Ha (table interpolation), mixed to 5e-7 Ha (1e-6 Ry)."""
    torch = _torch()
    wl = workload("si2_fcc")
    kp = wl.kpts[0]
    ref = np.linalg.eigvalsh(refmath.dense_h(wl.cell, kp))[:8]
    H = Hamiltonian(wl.cell)
    H.load_kpoint(0, kp)
    kin = torch.from_numpy(kp.kin).cuda()
    op = H.at_k(0)
    for prec, wdt, tol, etol in (("fp64", torch.complex128, 1e-9, 1e-7), ("mixed", torch.complex64, 2e-5, 5e-7)):
        X0 = eigensolver.random_start(12, kp.npw, kin, seed=5, device="cuda")
        res = eigensolver.chfsi(lambda V, HV: op(V, HV, prec=prec, overwrite=True), X0, kin, nconv=8, tol=tol,
                                maxiter=200, work_dtype=wdt, gram=H.overlap, degree=20,
                                eigh=eigensolver.eigh_host, lincomb=H.lincomb)
        assert res.converged == 8, prec
        assert np.abs(res.E[:8].cpu().numpy() - ref).max() < etol, prec
    H.close()
