"""Python mirror of the reference's H_psi module interface over libeffm_hip.so.

Reference: src/H_psi/hamiltonian.cppm:115-363.
  Hamiltonian.at_k(ikpt) -> Callable;  Callable(psi, hpsi)  (hpsi += H psi)
  Callable(psi, hpsi, n_bands=nb)      the declared batched form (:180-182)
  Hamiltonian.gradient().at_k(ikpt)(psi, GradHPsi(x, y, z))  (:218-263)
Semantics kept from the reference: results ACCUMULATE into the output
(src/H_psi/hamiltonian_callable.cpp:87-92); a None direction of GradHPsi is
skipped (hamiltonian.cppm:53-65).  Errors raise the classes of _lib that
subclass the reference's exception kinds (runtime / out_of_range /
domain_error).

psi / hpsi may be torch tensors on the GPU (complex128 for "fp64",
complex64 for "mixed"; shape [nb, npw] or [npw], last dim contiguous) or
numpy arrays on the host (staged through device buffers).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from . import _lib
from .synth import Cell as SynthCell, KPoint as SynthKPoint

PREC = {"fp64": _lib.EFFM_PREC_FP64, "mixed": _lib.EFFM_PREC_MIXED}


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _arr_info(x, prec: int, device: Optional[int] = None):
    """(pointer, nb, ld, host?, columns) of a band-major psi-like array."""
    want = np.complex128 if prec == _lib.EFFM_PREC_FP64 else np.complex64
    if _is_torch(x):
        import torch
        tdt = torch.complex128 if prec == _lib.EFFM_PREC_FP64 else torch.complex64
        if x.dtype != tdt:
            raise TypeError(f"expected {tdt}, got {x.dtype}")
        if not x.is_cuda:
            raise TypeError("torch tensors must live on the GPU; pass numpy arrays for host data")
        if device is not None and x.device.index != device:
            raise ValueError(f"tensor is on cuda:{x.device.index}, the Hamiltonian on cuda:{device}")
        if x.dim() not in (1, 2):
            raise ValueError("psi-like arrays are [npw] or [nb, npw]")
        x2 = x.view(1, -1) if x.dim() == 1 else x
        if x2.stride(-1) != 1:
            raise ValueError("last dimension must be contiguous")
        ld = x2.stride(0) if x2.shape[0] > 1 else x2.shape[1]
        return x.data_ptr(), x2.shape[0], ld, False, x2.shape[1]
    if isinstance(x, np.ndarray):
        if x.dtype != want:
            raise TypeError(f"expected {np.dtype(want)}, got {x.dtype}")
        if not x.flags.c_contiguous:
            raise ValueError("host arrays must be C-contiguous")
        if x.ndim not in (1, 2):
            raise ValueError("psi-like arrays are [npw] or [nb, npw]")
        x2 = x.reshape(1, -1) if x.ndim == 1 else x
        return x.ctypes.data, x2.shape[0], x2.shape[1], True, x2.shape[1]
    raise TypeError(f"unsupported array type {type(x)}")


def _bands(n_bands: Optional[int], dim: int, psi_info, out_infos) -> int:
    """Band count of a call, checked against every array before the library
    sees a pointer: each holds at least that many bands of at least dim()
    coefficients (the kernels read and write nb rows of npw elements)."""
    nb = psi_info[1]
    nbo = min((o[1] for o in out_infos), default=nb)
    if n_bands is None:
        if nbo < nb:
            raise ValueError(f"output holds {nbo} bands, psi {nb}")
    else:
        if not (0 <= n_bands <= min(nb, nbo)):
            raise ValueError(f"n_bands={n_bands} outside [0, {min(nb, nbo)}] (bands held by psi / outputs)")
        nb = n_bands
    for info, what in [(psi_info, "psi")] + [(o, "output") for o in out_infos]:
        if info[4] < dim:
            raise ValueError(f"{what} has {info[4]} coefficients per band, the k-point needs {dim}")
    return nb


@dataclass
class GradHPsi:
    """Output container of dH/dk|psi> (hamiltonian.cppm:66-70); None = skip."""
    x: Any = None
    y: Any = None
    z: Any = None


@dataclass
class HessHPsi:
    """Output container of d2H/dk_a dk_b |psi> (hamiltonian.cppm:84-91): the
    six unique components; None = skip."""
    xx: Any = None
    yy: Any = None
    zz: Any = None
    xy: Any = None
    xz: Any = None
    yz: Any = None


def shifted_kpoint(kp: SynthKPoint, A, dk_cart) -> SynthKPoint:
    """The same G set at k + dk (Cartesian): K + dk, |K|^2/2, k_frac + dk.A^T/2pi."""
    dk = np.asarray(dk_cart, dtype=np.float64)
    K = kp.K + dk[None, :]
    return SynthKPoint(kfrac=kp.kfrac + dk @ np.asarray(A, dtype=np.float64).T / (2.0 * np.pi),
                       K=np.ascontiguousarray(K), kin=0.5 * np.einsum("ij,ij->i", K, K), gidx=kp.gidx)


class Hamiltonian:
    """Owns the device-resident cell and k-points (hamiltonian.cppm:115)."""

    def __init__(self, cell: SynthCell, device: int = 0, hbm_budget: int = 0):
        self._lib = _lib.load()
        self._ctx = C.c_void_p()
        _lib.check(self._lib.effm_ctx_create(C.byref(self._ctx), device, hbm_budget), None,
                   "effm_ctx_create")
        self.device = device
        self._keep = []
        self.cell = cell
        self._upload_cell(cell)
        self._npw = {}
        self._kpts = {}          # host copies of the loaded k-points (the Hessian's shifted copies)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if self._ctx:
            self._lib.effm_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self):
        return self._ctx

    def _sync_stream(self):
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        self._lib.effm_set_stream(self._ctx, C.c_void_p(s))

    # -- data --------------------------------------------------------------
    def _upload_cell(self, cell: SynthCell):
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        sps = (_lib.Species * max(1, len(cell.species_list)))()
        for i, sp in enumerate(cell.species_list):
            bl = arr(sp.beta_l, np.int32)
            lam = arr(sp.lam, np.float64)
            bt = arr(sp.beta_tab, np.float64)
            dbt = arr(sp.dbeta_tab, np.float64) if sp.dbeta_tab is not None else None
            sps[i].lmax = sp.lmax
            sps[i].nbeta = sp.nbeta
            sps[i].beta_l = bl.ctypes.data_as(C.POINTER(C.c_int32))
            sps[i].lam = lam.ctypes.data_as(C.POINTER(C.c_double))
            sps[i].nq = sp.nq
            sps[i].dq = sp.dq
            sps[i].qmax = sp.qmax
            sps[i].beta_tab = bt.ctypes.data_as(C.POINTER(C.c_double))
            sps[i].dbeta_tab = dbt.ctypes.data_as(C.POINTER(C.c_double)) if dbt is not None else None
        vl = arr(cell.vloc, np.float64)
        spc = arr(cell.species, np.int32)
        tau = arr(cell.tau, np.float64)
        cc = _lib.Cell()
        cc.n1, cc.n2, cc.n3 = cell.n
        for i, v in enumerate(np.asarray(cell.A, dtype=np.float64).ravel()):
            cc.lattice[i] = v
        cc.vloc = vl.ctypes.data_as(C.POINTER(C.c_double))
        cc.natom = cell.natom
        cc.species = spc.ctypes.data_as(C.POINTER(C.c_int32))
        cc.tau = tau.ctypes.data_as(C.POINTER(C.c_double))
        cc.nspecies = len(cell.species_list)
        cc.sp = sps
        _lib.check(self._lib.effm_cell_upload(self._ctx, C.byref(cc)), self._ctx, "effm_cell_upload")

    def load_kpoint(self, ikpt: int, kp: SynthKPoint):
        """GKK::loadKPoint equivalent: make k-point ikpt resident."""
        K = np.ascontiguousarray(kp.K, dtype=np.float64)
        kin = np.ascontiguousarray(kp.kin, dtype=np.float64)
        gi = np.ascontiguousarray(kp.gidx, dtype=np.int32)
        kf = np.ascontiguousarray(kp.kfrac, dtype=np.float64)
        _lib.check(self._lib.effm_kpt_upload(self._ctx, ikpt, kp.npw, K.ctypes.data, kin.ctypes.data,
                                             gi.ctypes.data, kf.ctypes.data), self._ctx,
                   "effm_kpt_upload")
        self._npw[ikpt] = kp.npw
        self._kpts[ikpt] = kp

    def release_kpoint(self, ikpt: int):
        _lib.check(self._lib.effm_kpt_release(self._ctx, ikpt), self._ctx, "effm_kpt_release")
        self._npw.pop(ikpt, None)
        self._kpts.pop(ikpt, None)

    def mem_info(self):
        u, b = C.c_uint64(), C.c_uint64()
        self._lib.effm_mem_info(self._ctx, C.byref(u), C.byref(b))
        return u.value, b.value

    # -- operators -----------------------------------------------------------
    def at_k(self, ikpt: int) -> "Callable":
        return Callable(self, ikpt)

    def gradient(self) -> "Gradient":
        return Gradient(self)

    def hessian(self, delta: float = 1e-3) -> "Hessian":
        return Hessian(self, delta)

    def rayleigh(self, psi, hpsi, prec: str = "fp64"):
        """FP64 E_b = Re<psi_b|hpsi_b>/<psi_b|psi_b> (test_hpsi_eigen.cpp:49-56)."""
        p = PREC[prec]
        pi = _arr_info(psi, p, self.device)
        hi = _arr_info(hpsi, p, self.device)
        pp, nb, ld, host, npw = pi
        hp, nb2, ldh, host2, npw2 = hi
        if host != host2:
            raise TypeError("psi and hpsi must both be host or both be device arrays")
        if nb2 < nb or npw2 < npw:
            raise ValueError(f"hpsi [{nb2}, {npw2}] is smaller than psi [{nb}, {npw}]")
        if host:
            E = np.zeros(nb, dtype=np.float64)
            _lib.check(self._lib.effm_rayleigh(self._ctx, npw, nb, pp, ld, hp, ldh, p, E.ctypes.data,
                                               _lib.EFFM_F_HOST), self._ctx, "effm_rayleigh")
            return E
        import torch
        self._sync_stream()
        E = torch.zeros(nb, dtype=torch.float64, device=psi.device)
        _lib.check(self._lib.effm_rayleigh(self._ctx, npw, nb, pp, ld, hp, ldh, p, E.data_ptr(), 0),
                   self._ctx, "effm_rayleigh")
        return E

    def _batch(self, ikpts, psis, outs, p):
        """ctypes arrays of a multi-k batch: every tensor on the device, one
        band count, outputs with the leading dimension of their psi."""
        nk = len(ikpts)
        ks = (C.c_int * nk)(*ikpts)
        pp = (C.c_void_p * nk)()
        lds = (C.c_int64 * nk)()
        ops = [None if o is None else (C.c_void_p * nk)() for o in outs]
        nb = None
        for i in range(nk):
            a, nbi, ld, host, cols = _arr_info(psis[i], p, self.device)
            npw = self._npw.get(ikpts[i])
            if npw is not None and cols < npw:
                raise ValueError(f"psi of k-point {ikpts[i]} has {cols} coefficients per band, it needs {npw}")
            if host:
                raise ValueError("batched applies take device tensors")
            if nb is not None and nbi != nb:
                raise ValueError("every k-point needs the same band count")
            nb = nbi
            pp[i], lds[i] = a, ld
            for o, arr in zip(outs, ops):
                if o is None:
                    continue
                b, nbo, ldo, host_o, ocols = _arr_info(o[i], p, self.device)
                if host_o or ldo != ld or nbo < nbi or ocols < cols:
                    raise ValueError("outputs must be device tensors shaped like their psi")
                arr[i] = b
        return nk, ks, nb or 0, pp, lds, ops

    def apply_multi(self, ikpts, psis, hpsis, prec: str = "fp64", overwrite: bool = False,
                    graph: bool = False):
        """H|psi> for several k-points (ragged npw) in one batched call;
        device tensors [nb, >= npw(k)].  graph=True replays a captured
        hipGraph of the launches (same pointers and shapes as a previous
        call: the contents of psi may change, not the buffers)."""
        p = PREC[prec]
        nk, ks, nb, pp, lds, (hp,) = self._batch(ikpts, psis, [hpsis], p)
        self._sync_stream()
        flags = (_lib.EFFM_F_OVERWRITE if overwrite else 0) | (_lib.EFFM_F_GRAPH if graph else 0)
        _lib.check(self._lib.effm_hpsi_multi(self._ctx, nk, ks, nb, pp, lds, hp, p, flags),
                   self._ctx, "effm_hpsi_multi")
        return hpsis

    def gradient_multi(self, ikpts, psis, outs: "GradHPsi", prec: str = "fp64", overwrite: bool = False,
                       graph: bool = False):
        """dH/dk|psi> for several k-points in one batched call; outs.x / .y /
        .z are per-k lists of device tensors (None skips a direction)."""
        p = PREC[prec]
        nk, ks, nb, pp, lds, (ox, oy, oz) = self._batch(ikpts, psis, [outs.x, outs.y, outs.z], p)
        self._sync_stream()
        flags = (_lib.EFFM_F_OVERWRITE if overwrite else 0) | (_lib.EFFM_F_GRAPH if graph else 0)
        _lib.check(self._lib.effm_dhdk_multi(self._ctx, nk, ks, nb, pp, lds, ox, oy, oz, p, flags),
                   self._ctx, "effm_dhdk_multi")
        return outs

    def apply_with_gradient_multi(self, ikpts, psis, hpsis, outs: "GradHPsi", prec: str = "fp64",
                                  overwrite: bool = False, graph: bool = False):
        """H|psi> into hpsis and dH/dk|psi> into outs for several k-points,
        with one projection GEMM for both (the k.p consumer's pair); same
        results as apply_multi followed by gradient_multi."""
        p = PREC[prec]
        nk, ks, nb, pp, lds, (hp, ox, oy, oz) = self._batch(ikpts, psis, [hpsis, outs.x, outs.y, outs.z], p)
        self._sync_stream()
        flags = (_lib.EFFM_F_OVERWRITE if overwrite else 0) | (_lib.EFFM_F_GRAPH if graph else 0)
        _lib.check(self._lib.effm_hpsi_dhdk_multi(self._ctx, nk, ks, nb, pp, lds, hp, ox, oy, oz, p, flags),
                   self._ctx, "effm_hpsi_dhdk_multi")
        return hpsis, outs

    def lincomb(self, z, a, y, b, x, c):
        """z <- a z + b y + c x in place for same-shaped contiguous device
        tensors (complex128 or complex64), one pass on the device."""
        import torch
        if not (z.dtype == y.dtype == x.dtype) or z.dtype not in (torch.complex128, torch.complex64):
            raise TypeError("lincomb takes three complex128 or three complex64 tensors")
        if not (z.shape == y.shape == x.shape) or not (z.is_contiguous() and y.is_contiguous()
                                                        and x.is_contiguous()):
            raise ValueError("lincomb takes same-shaped contiguous tensors")
        p = _lib.EFFM_PREC_FP64 if z.dtype == torch.complex128 else _lib.EFFM_PREC_MIXED
        self._sync_stream()
        _lib.check(self._lib.effm_lincomb(self._ctx, z.numel(), float(a), z.data_ptr(), float(b), y.data_ptr(),
                                          float(c), x.data_ptr(), p, 0), self._ctx, "effm_lincomb")
        return z

    def overlap(self, A, B):
        """C = conj(A) @ B.T for device tensors A [m, k], B [n, k] (complex128
        -> FP64 path, complex64 -> mixed), split-K sums added in FP64; C has
        A's dtype."""
        import torch
        if A.dtype != B.dtype or A.dtype not in (torch.complex128, torch.complex64):
            raise TypeError("overlap takes two complex128 or two complex64 tensors")
        A = A.contiguous()
        B = B.contiguous()
        m, k = A.shape
        n = B.shape[0]
        if B.shape[1] != k:
            raise ValueError("A and B need the same number of columns")
        Cm = torch.empty((m, n), dtype=A.dtype, device=A.device)
        p = _lib.EFFM_PREC_FP64 if A.dtype == torch.complex128 else _lib.EFFM_PREC_MIXED
        self._sync_stream()
        _lib.check(self._lib.effm_overlap(self._ctx, m, n, k, A.data_ptr(), max(k, 1), B.data_ptr(), max(k, 1),
                                          Cm.data_ptr(), max(n, 1), p, 0), self._ctx, "effm_overlap")
        return Cm

    def kp_mass(self, E, V, bands, d2h, degen_tol: float = 1e-6):
        """FP64 k.p sum over states on the device (effm_kp_mass): for each n
        in `bands`, (1/m*)_ab = sym(d2h_ab) + 2 Re sum_m V_a[n][m] V_b[m][n] /
        (E_n - E_m) over |E_n - E_m| > degen_tol, m != n.  E: float64 [nb],
        V: three complex128 [nb, nb] (V_a[m][l] = <m|dH/dk_a|l>), d2h:
        float64 [len(bands), 3, 3], all on this device; returns a float64
        [len(bands), 3, 3] device tensor."""
        import torch
        dev = torch.device("cuda", self.device)
        E = E.to(device=dev, dtype=torch.float64).contiguous()
        nb = E.numel()
        Vs = [v.to(device=dev, dtype=torch.complex128).contiguous() for v in V]
        if len(Vs) != 3 or any(tuple(v.shape) != (nb, nb) for v in Vs):
            raise ValueError(f"V must be three [{nb}, {nb}] matrices")
        bands = [int(n) for n in bands]
        d2 = d2h.to(device=dev, dtype=torch.float64).contiguous()
        if tuple(d2.shape) != (len(bands), 3, 3):
            raise ValueError("d2h must be [len(bands), 3, 3]")
        out = torch.empty((len(bands), 3, 3), dtype=torch.float64, device=dev)
        vp = (C.c_void_p * 3)(*[v.data_ptr() for v in Vs])
        bs = (C.c_int * max(1, len(bands)))(*bands)
        self._sync_stream()
        _lib.check(self._lib.effm_kp_mass(self._ctx, nb, E.data_ptr(), vp, nb, len(bands), bs, d2.data_ptr(),
                                          float(degen_tol), out.data_ptr(), 0), self._ctx, "effm_kp_mass")
        return out

    def autotune(self, ikpt: int, n_bands: int, prec: str = "fp64"):
        """Time the candidate launch shapes at this k-point and keep the
        fastest for `prec`; returns (bands per pencil workgroup, split-K
        workgroup target)."""
        pb, gw = C.c_int(), C.c_int()
        self._sync_stream()
        _lib.check(self._lib.effm_autotune(self._ctx, ikpt, n_bands, PREC[prec], C.byref(pb), C.byref(gw)),
                   self._ctx, "effm_autotune")
        return pb.value, gw.value

    SHAPE_KEYS = ("pencil_bands", "pencil_z", "gemm_wgs", "apply_tile")

    def shapes(self, prec: str = "fp64") -> dict:
        """The launch shapes in effect for `prec` (effm_get_shapes)."""
        s = _lib.Shapes()
        _lib.check(self._lib.effm_get_shapes(self._ctx, PREC[prec], C.byref(s)), self._ctx, "effm_get_shapes")
        return {k: getattr(s, k) for k in self.SHAPE_KEYS}

    def set_shapes(self, prec: str = "fp64", **shapes):
        """Set some of the launch shapes of `prec`; the others keep the
        values in effect.  A value outside its candidate list raises
        (effm_set_shapes)."""
        unknown = set(shapes) - set(self.SHAPE_KEYS)
        if unknown:
            raise TypeError(f"unknown launch shapes: {sorted(unknown)}")
        cur = self.shapes(prec)
        cur.update(shapes)
        self._sync_stream()
        _lib.check(self._lib.effm_set_shapes(self._ctx, PREC[prec], C.byref(_lib.Shapes(**cur))), self._ctx,
                   "effm_set_shapes")

    def set_timing(self, on: bool):
        self._lib.effm_set_timing(self._ctx, 1 if on else 0)

    def timing(self, name: str):
        ms, cnt = C.c_double(), C.c_int64()
        self._lib.effm_get_timing(self._ctx, name.encode(), C.byref(ms), C.byref(cnt))
        return ms.value, cnt.value

    def reset_timing(self):
        self._lib.effm_reset_timing(self._ctx)


class Callable:
    """H|psi> at fixed k (hamiltonian.cppm:173-204)."""

    def __init__(self, parent: Hamiltonian, ikpt: int):
        self._h = parent
        self._ikpt = ikpt
        self.dim()   # raises out_of_range for an unknown k-point

    def dim(self) -> int:
        return _lib.check(self._h._lib.effm_kpt_dim(self._h._ctx, self._ikpt), self._h._ctx,
                          "Callable::dim")

    def set_ikpt(self, ikpt: int):
        self._ikpt = ikpt
        self.dim()

    def __call__(self, psi, hpsi, n_bands: Optional[int] = None, prec: str = "fp64",
                 overwrite: bool = False, graph: bool = False):
        p = PREC[prec]
        dev = self._h.device
        pi = _arr_info(psi, p, dev)
        hi = _arr_info(hpsi, p, dev)
        pp, _, ld, host, _ = pi
        hp, _, ldo, host_o, _ = hi
        if host != host_o:
            raise TypeError("psi and hpsi must both be host or both be device arrays")
        nb = _bands(n_bands, self.dim(), pi, [hi])
        flags = ((_lib.EFFM_F_OVERWRITE if overwrite else 0) | (_lib.EFFM_F_HOST if host else 0)
                 | (_lib.EFFM_F_GRAPH if graph else 0))
        if not host:
            self._h._sync_stream()
        _lib.check(self._h._lib.effm_hpsi(self._h._ctx, self._ikpt, nb, pp, ld, hp, ldo, p, flags),
                   self._h._ctx, "H|psi>")
        return hpsi


class Gradient:
    def __init__(self, parent: Hamiltonian):
        self._h = parent

    def at_k(self, ikpt: int) -> "GradientCallable":
        return GradientCallable(self._h, ikpt)


class GradientCallable:
    """dH/dk_alpha |psi> at fixed k (hamiltonian.cppm:222-254)."""

    def __init__(self, parent: Hamiltonian, ikpt: int):
        self._h = parent
        self._ikpt = ikpt
        self.dim()

    def dim(self) -> int:
        return _lib.check(self._h._lib.effm_kpt_dim(self._h._ctx, self._ikpt), self._h._ctx,
                          "Gradient::Callable::dim")

    def __call__(self, psi, out: GradHPsi, prec: str = "fp64", overwrite: bool = False,
                 n_bands: Optional[int] = None):
        p = PREC[prec]
        dev = self._h.device
        pi = _arr_info(psi, p, dev)
        pp, _, ld, host, _ = pi
        ptrs = []
        infos = []
        ldo = None
        for o in (out.x, out.y, out.z):
            if o is None:
                ptrs.append(None)
                continue
            oi = _arr_info(o, p, dev)
            op, _, l2, host_o, _ = oi
            if host_o != host:
                raise TypeError("psi and outputs must both be host or both be device arrays")
            if ldo is not None and l2 != ldo:
                raise ValueError("all outputs need the same leading dimension")
            ldo = l2
            ptrs.append(op)
            infos.append(oi)
        if ldo is None:
            return out
        nb = _bands(n_bands, self.dim(), pi, infos)
        flags = (_lib.EFFM_F_OVERWRITE if overwrite else 0) | (_lib.EFFM_F_HOST if host else 0)
        if not host:
            self._h._sync_stream()
        _lib.check(self._h._lib.effm_dhdk(self._h._ctx, self._ikpt, nb, pp, ld, ptrs[0], ptrs[1], ptrs[2],
                                          ldo, p, flags), self._h._ctx, "dH/dk|psi>")
        return out


class Hessian:
    def __init__(self, parent: Hamiltonian, delta: float = 1e-3):
        self._h = parent
        self._delta = delta

    def at_k(self, ikpt: int) -> "HessianCallable":
        return HessianCallable(self._h, ikpt, self._delta)


class HessianCallable:
    """d2H/dk_a dk_b |psi> at fixed k, the operator the reference declares
    (hamiltonian.cppm:267-306) and leaves empty (hamiltonian_hessian.cpp:48-50).
    d2H/dk_a dk_b = delta_ab (kinetic) + d2V_NL/dk_a dk_b; here the whole of
    it is the central difference over k_b of the analytic dH/dk_a at
    k +- delta e_b (Cartesian, bohr^-1) on the same G set, which is exact for
    the kinetic part (dH/dk_a = K_a there) and O(delta^2) for the nonlocal
    part; xy, xz, yz take the difference of dH/dk_x, dH/dk_x, dH/dk_y along
    y, z, z.  The shifted k-points go through one scratch slot past the
    highest loaded one, released afterwards.  Results accumulate into the
    outputs (overwrite=True replaces them); a None component is skipped.
    psi / outputs: device tensors or host arrays, as for Callable; the
    difference is taken in FP64."""

    # shift direction b -> the components it gives (name, a)
    _PLAN = ((0, (("xx", 0),)), (1, (("xy", 0), ("yy", 1))), (2, (("xz", 0), ("yz", 1), ("zz", 2))))

    def __init__(self, parent: Hamiltonian, ikpt: int, delta: float):
        self._h = parent
        self._ikpt = ikpt
        self._delta = float(delta)
        self.dim()

    def dim(self) -> int:
        return _lib.check(self._h._lib.effm_kpt_dim(self._h._ctx, self._ikpt), self._h._ctx,
                          "Hessian::Callable::dim")

    def set_ikpt(self, ikpt: int):
        self._ikpt = ikpt
        self.dim()

    def __call__(self, psi, out: HessHPsi, prec: str = "fp64", overwrite: bool = False,
                 n_bands: Optional[int] = None):
        import torch
        p = PREC[prec]
        dev = self._h.device
        pi = _arr_info(psi, p, dev)
        names = [n for n in ("xx", "yy", "zz", "xy", "xz", "yz") if getattr(out, n) is not None]
        infos = []
        for n in names:
            oi = _arr_info(getattr(out, n), p, dev)
            if oi[3] != pi[3]:
                raise TypeError("psi and outputs must both be host or both be device arrays")
            infos.append(oi)
        if not names:
            return out
        nb = _bands(n_bands, self.dim(), pi, infos)
        kp = self._h._kpts.get(self._ikpt)
        if kp is None:
            raise _lib.EffmOutOfRange(f"k-point {self._ikpt} was not loaded through this Hamiltonian")
        npw = kp.npw
        cdt = torch.complex128 if prec == "fp64" else torch.complex64
        x = torch.as_tensor(psi, device=f"cuda:{dev}").reshape(-1, psi.shape[-1])[:nb, :npw]
        x = x.to(cdt).contiguous()
        slot = max(self._h._npw.keys()) + 1
        res = {}
        try:
            for b, comps in self._PLAN:
                want = [(n, a) for n, a in comps if n in names]
                if not want:
                    continue
                D = {}
                for sgn in (+1.0, -1.0):
                    dk = np.zeros(3)
                    dk[b] = sgn * self._delta
                    self._h.load_kpoint(slot, shifted_kpoint(kp, self._h.cell.A, dk))
                    g = [torch.zeros_like(x) if any(a == d for _, a in want) else None for d in range(3)]
                    self._h.gradient().at_k(slot)(x, GradHPsi(*g), prec=prec, overwrite=True)
                    D[sgn] = g
                for n, a in want:
                    res[n] = (D[1.0][a].to(torch.complex128) - D[-1.0][a].to(torch.complex128)) / (2 * self._delta)
        finally:
            if slot in self._h._npw:
                self._h.release_kpoint(slot)
        for n in names:
            o = getattr(out, n)
            v = res[n]
            if _is_torch(o):
                tgt = o.reshape(-1, o.shape[-1])[:nb, :npw]
                if overwrite:
                    tgt.copy_(v.to(o.dtype))
                else:
                    tgt.add_(v.to(o.dtype))
            else:
                tgt = o.reshape(-1, o.shape[-1])[:nb, :npw]
                vh = v.cpu().numpy().astype(o.dtype)
                if overwrite:
                    tgt[...] = vh
                else:
                    tgt += vh
        return out
