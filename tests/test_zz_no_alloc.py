"""dis_coil_gram_noalloc (the allocation-free variant of dis_coil_gram,
whose block partials live in a static per-GPU device array) against
dis_coil_gram through the C-ABI: two calls in a row, the second on a
smaller input (fewer voxels and coils).  The static variant runs the VALU
Gram kernel (FP32 products flushed into FP64 every 256 voxels, at most 256
blocks so its partials fit the static array) and dis_coil_gram the
split-bf16 MFMA kernel; both are pinned to rel-L2 1e-6 against FP64 in
test_coilcomp.py, so they agree to 2e-6 relative."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _gram_noalloc(ctx, S, nv, C, dev):
    P = torch.zeros(2 * C * C, dtype=torch.float64, device=dev)
    ctx._prep()
    st = ctx.lib.dis_coil_gram_noalloc(ctx.h, ctypes.c_void_p(S.data_ptr()), nv, C, ctypes.c_void_p(P.data_ptr()))
    assert st == 0, ctx.lib.dis_last_error(ctx.h)
    assert ctx.lib.dis_sync(ctx.h) == 0
    p = P.cpu().numpy()
    return (p[0::2] + 1j * p[1::2]).reshape(C, C, order="F")


def _gram(ctx, S, nv, C):
    P = np.zeros(2 * C * C, np.float64)
    ctx._prep()
    st = ctx.lib.dis_coil_gram(ctx.h, ctypes.c_void_p(S.data_ptr()), nv, C,
                               P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert st == 0, ctx.lib.dis_last_error(ctx.h)
    return (P[0::2] + 1j * P[1::2]).reshape(C, C, order="F")


def test_noalloc_gram_matches_allocating_gram(dev):
    import disorder_amd as D
    ctx = D.default_context(0)
    g = torch.Generator(device=dev).manual_seed(5)
    for nv, C in ((3_000_017, 40), (70_001, 12)):  # second call: smaller input
        S = torch.randn((C, nv), dtype=torch.complex64, device=dev, generator=g)  # row c = coil c: stride nv
        a = _gram_noalloc(ctx, S, nv, C, dev)
        b = _gram(ctx, S, nv, C)
        assert np.linalg.norm(a - b) <= 2e-6 * np.linalg.norm(b), (nv, C, np.linalg.norm(a - b) / np.linalg.norm(b))
        assert np.allclose(a, a.conj().T, rtol=0, atol=1e-7 * np.abs(a).max())  # Hermitian
        assert np.all(np.abs(np.diag(a).imag) <= 1e-6 * np.abs(np.diag(a).real))


def test_noalloc_gram_rejects_more_than_64_coils(dev):
    import disorder_amd as D
    ctx = D.default_context(0)
    S = torch.zeros((65, 64), dtype=torch.complex64, device=dev)
    P = torch.zeros(2 * 65 * 65, dtype=torch.float64, device=dev)
    st = ctx.lib.dis_coil_gram_noalloc(ctx.h, ctypes.c_void_p(S.data_ptr()), 64, 65, ctypes.c_void_p(P.data_ptr()))
    assert st == -2 and b"64" in ctx.lib.dis_last_error(ctx.h)
