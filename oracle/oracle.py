"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")


class OrcSpecies(C.Structure):
    _fields_ = [
        ("lmax", C.c_int), ("nbeta", C.c_int), ("beta_l", C.POINTER(C.c_int32)),
        ("lam", C.POINTER(C.c_double)), ("nq", C.c_int), ("dq", C.c_double), ("qmax", C.c_double),
        ("beta_tab", C.POINTER(C.c_double)), ("dbeta_tab", C.POINTER(C.c_double)),
    ]


class OrcCell(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("vloc", C.POINTER(C.c_double)),
        ("omega", C.c_double), ("natom", C.c_int), ("species", C.POINTER(C.c_int)),
        ("tau", C.POINTER(C.c_double)), ("nspecies", C.c_int), ("sp", C.POINTER(OrcSpecies)),
    ]


class OrcKpt(C.Structure):
    _fields_ = [
        ("npw", C.c_int), ("K", C.POINTER(C.c_double)), ("kin", C.POINTER(C.c_double)),
        ("gidx", C.POINTER(C.c_int32)), ("kfrac", C.c_double * 3),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    lib.orc_fft1d.argtypes = [dp, C.c_int, C.c_int]
    lib.orc_fft3d.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.orc_interp.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp]
    lib.orc_ylm.argtypes = [C.c_double, C.c_double, C.c_int, dp, dp, dp]
    lib.orc_sfac.argtypes = [dp, C.POINTER(C.c_int), dp, C.c_int, C.c_int, C.c_int, dp]
    for f in ("orc_hpsi", "orc_vnl"):
        getattr(lib, f).argtypes = [C.POINTER(OrcCell), C.POINTER(OrcKpt), C.c_int, dp, dp]
        getattr(lib, f).restype = C.c_int
    lib.orc_hpsi_parallel.argtypes = [C.POINTER(OrcCell), C.POINTER(OrcKpt), C.c_int, dp, dp, C.c_int]
    lib.orc_dhdk.argtypes = [C.POINTER(OrcCell), C.POINTER(OrcKpt), C.c_int, dp, dp, dp, dp]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class OracleCell:
    """Keeps the numpy buffers a C orc_cell points into alive."""

    def __init__(self, cell, drop_dbeta=False):
        self.keep = []
        sps = (OrcSpecies * max(1, len(cell.species_list)))()
        for i, sp in enumerate(cell.species_list):
            bl = np.ascontiguousarray(sp.beta_l, np.int32)
            lam = np.ascontiguousarray(sp.lam, np.float64)
            bt = np.ascontiguousarray(sp.beta_tab, np.float64)
            dbt = None if (drop_dbeta or sp.dbeta_tab is None) else np.ascontiguousarray(sp.dbeta_tab, np.float64)
            self.keep += [bl, lam, bt, dbt]
            sps[i].lmax, sps[i].nbeta = sp.lmax, sp.nbeta
            sps[i].beta_l = bl.ctypes.data_as(C.POINTER(C.c_int32))
            sps[i].lam = _dp(lam)
            sps[i].nq, sps[i].dq, sps[i].qmax = sp.nq, sp.dq, sp.qmax
            sps[i].beta_tab = _dp(bt)
            sps[i].dbeta_tab = _dp(dbt)
        vl = np.ascontiguousarray(cell.vloc, np.float64)
        spc = np.ascontiguousarray(cell.species, np.int32)
        tau = np.ascontiguousarray(cell.tau, np.float64)
        self.keep += [vl, spc, tau, sps]
        c = OrcCell()
        c.n1, c.n2, c.n3 = cell.n
        c.vloc = _dp(vl)
        c.omega = cell.omega
        c.natom = cell.natom
        c.species = spc.ctypes.data_as(C.POINTER(C.c_int))
        c.tau = _dp(tau)
        c.nspecies = len(cell.species_list)
        c.sp = sps
        self.c = c


class OracleKpt:
    def __init__(self, kp):
        self.K = np.ascontiguousarray(kp.K, np.float64)
        self.kin = np.ascontiguousarray(kp.kin, np.float64)
        self.gidx = np.ascontiguousarray(kp.gidx, np.int32)
        k = OrcKpt()
        k.npw = kp.npw
        k.K = _dp(self.K)
        k.kin = _dp(self.kin)
        k.gidx = self.gidx.ctypes.data_as(C.POINTER(C.c_int32))
        for d in range(3):
            k.kfrac[d] = float(kp.kfrac[d])
        self.c = k


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def hpsi(cell, kp, psi, hpsi_in=None, nonlocal_only=False, nthreads=0):
    """hpsi_in + H psi (accumulate, as the reference)."""
    lib = load()
    oc, ok = OracleCell(cell), OracleKpt(kp)
    psi = np.ascontiguousarray(np.atleast_2d(psi), np.complex128)
    out = np.zeros_like(psi) if hpsi_in is None else np.array(hpsi_in, dtype=np.complex128, copy=True)
    pp = psi.view(np.float64)
    po = out.view(np.float64)
    if nthreads:
        rc = lib.orc_hpsi_parallel(C.byref(oc.c), C.byref(ok.c), psi.shape[0], _dp(pp), _dp(po), nthreads)
    elif nonlocal_only:
        rc = lib.orc_vnl(C.byref(oc.c), C.byref(ok.c), psi.shape[0], _dp(pp), _dp(po))
    else:
        rc = lib.orc_hpsi(C.byref(oc.c), C.byref(ok.c), psi.shape[0], _dp(pp), _dp(po))
    _check(rc, "hpsi")
    return out


def dhdk(cell, kp, psi, dirs=(True, True, True)):
    lib = load()
    oc, ok = OracleCell(cell), OracleKpt(kp)
    psi = np.ascontiguousarray(np.atleast_2d(psi), np.complex128)
    outs = [np.zeros_like(psi) if d else None for d in dirs]
    ptrs = [_dp(o.view(np.float64)) if o is not None else None for o in outs]
    rc = lib.orc_dhdk(C.byref(oc.c), C.byref(ok.c), psi.shape[0], _dp(psi.view(np.float64)), *ptrs)
    _check(rc, "dhdk")
    return outs


def fft1d(x, forward):
    lib = load()
    a = np.ascontiguousarray(x, np.complex128).copy()
    lib.orc_fft1d(_dp(a.view(np.float64)), a.size, 1 if forward else 0)
    return a


def fft3d(x, forward):
    lib = load()
    a = np.ascontiguousarray(x, np.complex128).copy()
    n1, n2, n3 = a.shape
    lib.orc_fft3d(_dp(a.view(np.float64)), n1, n2, n3, 1 if forward else 0)
    return a


def interp(tab, dq, qmax, norm, q):
    lib = load()
    t = np.ascontiguousarray(tab, np.float64)
    out = C.c_double()
    rc = lib.orc_interp(_dp(t), t.size, dq, qmax, norm, q, C.byref(out))
    return rc, out.value


def ylm(theta, phi, lmax):
    lib = load()
    n = (lmax + 1) ** 2
    y = np.zeros(n)
    dt = np.zeros(n)
    dp = np.zeros(n)
    lib.orc_ylm(theta, phi, lmax, _dp(y), _dp(dt), _dp(dp))
    return y, dt, dp


def sfac(tau, g, kfrac, n):
    lib = load()
    t = np.ascontiguousarray(tau, np.float64)
    gg = (C.c_int * 3)(*[int(v) for v in g])
    kf = np.ascontiguousarray(kfrac, np.float64)
    out = np.zeros(2)
    lib.orc_sfac(_dp(t), gg, _dp(kf), n[0], n[1], n[2], _dp(out))
    return complex(out[0], out[1])
