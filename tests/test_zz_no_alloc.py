"""dis_coil_gram_noalloc (the allocation-free variant of dis_coil_gram,
whose block partials live in a static per-GPU device array) against
dis_coil_gram through the C-ABI: two calls in a row, the second on a
smaller input (fewer voxels and coils).  The kernels are shared: with up to
256 x 4096 voxels both entry points cut the voxels into the same blocks and
agree bit for bit; above that the static variant uses 256 larger blocks,
whose FP32 partial sums (256 products each before the FP64 flush) round
differently, so the bound is the Gram's FP32-product accuracy (1e-7 of the
largest entry; the Gram itself is pinned to 1e-6 against FP64 in
test_coilcomp.py)."""
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
        if nv <= 256 * 4096:
            assert np.array_equal(a, b), (nv, C, np.abs(a - b).max())
        else:
            assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max(), (nv, C, np.abs(a - b).max())
        assert np.allclose(a, a.conj().T, rtol=0, atol=1e-7 * np.abs(a).max())  # Hermitian
        assert np.all(np.abs(np.diag(a).imag) <= 1e-6 * np.abs(np.diag(a).real))


def test_noalloc_gram_rejects_more_than_64_coils(dev):
    import disorder_amd as D
    ctx = D.default_context(0)
    S = torch.zeros((65, 64), dtype=torch.complex64, device=dev)
    P = torch.zeros(2 * 65 * 65, dtype=torch.float64, device=dev)
    st = ctx.lib.dis_coil_gram_noalloc(ctx.h, ctypes.c_void_p(S.data_ptr()), 64, 65, ctypes.c_void_p(P.data_ptr()))
    assert st == -2 and b"64" in ctx.lib.dis_last_error(ctx.h)
