"""effm_overlap_noalloc (static device scratch, no allocation) against
effm_overlap through the C-ABI: bit-identical results on two calls in a row,
the second smaller, in both precisions; the context allocates nothing; a
size whose partial sums exceed the 32 MiB scratch is refused before any
launch."""
import ctypes as C
import sys

import numpy as np
import pytest

from effmass_amd import _lib
from effmass_amd.hamiltonian import Hamiltonian  # noqa: F401  (loads the library the same way)

pytestmark = pytest.mark.gpu


def _ctx(lib):
    ctx = C.c_void_p()
    _lib.check(lib.effm_ctx_create(C.byref(ctx), 0, 0), None, "effm_ctx_create")
    return ctx


def _used(lib, ctx):
    u, b = C.c_uint64(), C.c_uint64()
    lib.effm_mem_info(ctx, C.byref(u), C.byref(b))
    return u.value


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_overlap_equals_overlap(prec):
    import torch
    assert torch.cuda.is_available()
    lib = _lib.load()
    ctx = _ctx(lib)
    p = _lib.EFFM_PREC_FP64 if prec == "fp64" else _lib.EFFM_PREC_MIXED
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    g = torch.Generator(device="cuda").manual_seed(5)
    s = torch.cuda.current_stream().cuda_stream
    lib.effm_set_stream(ctx, C.c_void_p(s))
    try:
        # two no-allocation calls in a row on one context, the second smaller
        # (m, n, k); then effm_overlap on the same operands
        shapes = ((256, 200, 36945), (37, 101, 777))
        ops, outs = [], []
        for (m, n, k) in shapes:
            print(f"{prec}: effm_overlap_noalloc m {m} n {n} k {k}", flush=True)
            A = torch.randn(m, k, dtype=dt, device="cuda", generator=g)
            B = torch.randn(n, k, dtype=dt, device="cuda", generator=g)
            C1 = torch.full((m, n), 7.0, dtype=dt, device="cuda")
            _lib.check(lib.effm_overlap_noalloc(ctx, m, n, k, A.data_ptr(), k, B.data_ptr(), k, C1.data_ptr(), n,
                                                p, 0), ctx, "effm_overlap_noalloc")
            torch.cuda.synchronize()
            assert _used(lib, ctx) == 0                    # the context allocated nothing
            ops.append((A, B))
            outs.append(C1)
        for (m, n, k), (A, B), C1 in zip(shapes, ops, outs):
            C2 = torch.zeros((m, n), dtype=dt, device="cuda")
            _lib.check(lib.effm_overlap(ctx, m, n, k, A.data_ptr(), k, B.data_ptr(), k, C2.data_ptr(), n, p, 0),
                       ctx, "effm_overlap")
            torch.cuda.synchronize()
            assert torch.equal(C1, C2), (m, n, k)
            ref = A.conj().to(torch.complex128) @ B.to(torch.complex128).T
            tol = 1e-12 if prec == "fp64" else 2e-5
            assert ((C1.to(torch.complex128) - ref).abs().max() / ref.abs().max()).item() < tol
        # partial sums beyond the scratch (one split's m x n block alone is
        # 72 MB mixed, 144 MB FP64) are refused on the host, nothing launched
        m = n = 3000
        rc = lib.effm_overlap_noalloc(ctx, m, n, 64, 1, 64, 1, 64, 1, n, p, 0)
        assert rc == _lib.EFFM_ERR_INVALID
        assert b"static scratch" in lib.effm_last_error(ctx)
    finally:
        lib.effm_ctx_destroy(ctx)
    sys.stdout.flush()
