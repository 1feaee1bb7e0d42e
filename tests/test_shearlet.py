"""Shearlet-regularised final reconstruction (Reconstruction/CSsolver.m,
regularize.m:27-50, traceEstimation.m, Shearlets/buildShearlet.m).

CPU: helpers against their definitions; closed properties of the shearlet
planes; product planes and band table against the oracle's literal
construction; the regulariser's self-adjointness, positivity and unit-weight
closed form; trace estimates against their unit-weight closed forms.
GPU: IRLS weights, the regulariser, both trace probes and the whole
CSsolver (lambdas and image) against the FP64 oracle."""
import math

import numpy as np
import pytest

from disorder_amd import shearlet as SH
from oracle import shearlet_oracle as SO
from tests.helpers import crandn, rel_l2

N3 = (40, 36, 35)


def test_helpers_match_definitions():
    x = np.arange(1.0, 5.0)[None, :]
    assert SO.upsampling(x, 1, 3).tolist() == [[1, 0, 0, 2, 0, 0, 3, 0, 0, 4]]
    # centred pad: zeroF = ceil((8+1)/2) = 5 (1-based); 4 values span 5-2 .. 5+1
    assert SO.resample_fo2(x, (1, 8)).tolist() == [[0, 0, 1, 2, 3, 4, 0, 0]]
    # crop 9 -> 4: zeroF = 5, kept 1-based 3..6
    assert SO.resample_fo2(np.arange(1.0, 10.0)[None, :], (1, 4)).tolist() == [[3, 4, 5, 6]]
    rng = np.random.default_rng(0)
    A = rng.normal(size=(7, 10))
    for k in (1, -2):
        r = np.arange(1, 8) - (7 // 2 + 1)
        ref = np.stack([np.roll(A[i], -k * r[i]) for i in range(7)])
        assert np.abs(SO.shearing(A, k).real - ref).max() < 1e-12
    for Na, Nb in ((40, 36), (33, 48)):
        z = rng.normal(size=(Na, Nb))
        assert np.allclose(SH._centre(z, (Na + 5, Nb - 4)), SO.resample_fo2(z, (Na + 5, Nb - 4)))
        assert np.allclose(SH._zero_insert(z, 1, 2), SO.upsampling(z, 1, 2))


def _neg(A):
    return np.roll(np.flip(np.roll(np.flip(A, 0), 1, 0), 1), 1, 1)


def test_plane_properties():
    P = SO.shearlets_2d(40, 36)
    Q = SO.shearlets_2d(36, 40)
    assert abs(P[0, 0, 10] - 1.0) < 1e-12
    assert np.abs(P[..., 10] - np.conj(_neg(P[..., 10]))).max() < 1e-12
    assert np.abs(P[0, 0, :10]).max() < 1e-12
    for i in range(5):
        assert np.array_equal(P[:, :, 5 + i], Q[:, :, 4 - i].T)
    S3 = SO.shearlets_3d(N3)
    assert abs(S3[0, 0, 0, -1] - 1.0) < 1e-12 and np.abs(S3[0, 0, 0, :-1]).max() < 1e-12
    with pytest.raises(ValueError):
        SO.shearlets_2d(32, 40)
    with pytest.raises(ValueError):
        SH.plane_filters(40, 32)


@pytest.mark.parametrize("Na,Nb", [(40, 36), (36, 40), (35, 48), (33, 33)])
def test_product_planes_match_oracle(Na, Nb):
    assert np.abs(SH.plane_filters(Na, Nb) - SO.shearlets_2d(Na, Nb)).max() < 1e-12


def test_band_table_matches_oracle_bands():
    S3 = SO.shearlets_3d(N3)
    n1, n2, n3 = N3
    P = [SH.plane_filters(n1, n2), SH.plane_filters(n1, n3), SH.plane_filters(n3, n2)]
    k1, k2, k3 = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")

    def pv(p, i):
        return P[0][k1, k2, i] if p == 0 else P[1][k1, k3, i] if p == 1 else P[2][k3, k2, i]
    bt = SH.band_table()
    assert len(bt) == 50 == S3.shape[-1]
    for w, (a, b, c, d) in enumerate(bt):
        assert np.abs(pv(a, b) * pv(c, d) - S3[..., w]).max() < 1e-12, w


def test_regulariser_properties():
    rng = np.random.default_rng(1)
    S = SO.shearlets_3d(N3)
    x, z = crandn(rng, N3), crandn(rng, N3)
    We = SO.shearlet_weights(crandn(rng, N3), S)
    Rx, Rz = SO.regularize_shearlet(x, S, We, 2.0), SO.regularize_shearlet(z, S, We, 2.0)
    assert abs(np.vdot(z, Rx) - np.vdot(Rz, x)) < 1e-10 * abs(np.vdot(z, Rx))
    assert np.vdot(x, Rx).real > 0 and abs(np.vdot(x, Rx).imag) < 1e-10 * abs(np.vdot(x, Rx))
    ones = np.ones(S.shape[:3] + (S.shape[3],))
    ref = np.fft.ifftn(np.sum(np.abs(S) ** 2, axis=-1) * np.fft.fftn(x))
    assert rel_l2(SO.regularize_shearlet(x, S, ones, 2.0), ref) < 1e-12


def test_traces_converge_to_closed_forms():
    rng = np.random.default_rng(2)
    S = SO.shearlets_3d(N3)
    ones = np.ones(S.shape)
    tr = SO.trace_estimation(lambda pr: SO.trace_samples_regularizer(pr, S, ones), N3, rng, 1e-2)
    assert abs(tr / np.sum(np.abs(S) ** 2) - 1) < 0.03
    Sm = crandn(rng, N3 + (3,))
    W = np.ones(N3[:2])
    tf = SO.trace_estimation(lambda pr: SO.trace_samples_fidelity(pr, Sm, W), N3, rng, 1e-2)
    assert abs(tf / np.sum(np.abs(Sm) ** 2) - 1) < 0.03


def _to_dev(a, dev):
    import disorder_amd as D
    import torch
    return D.colmajor(torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64))).to(dev))


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [N3, (64, 40, 64), (64, 64, 32 + 32), (144, 36, 48), (48, 36, 144)])
def test_gpu_weights_regulariser_and_traces_match_oracle(dev, grid):
    """(40,36,35): every axis on the LDS engine, regulariser without the fused
    middle pass; (64,40,64): axes 0 and 2 on the register path, axis 1 on the
    LDS engine, fused middle pass; (64,64,64): register path throughout;
    (144,36,48) and (48,36,144): the odd radix-9 register path on the
    contiguous axis and in the fused middle pass."""
    import torch
    rng = np.random.default_rng(3)
    x = crandn(rng, grid)
    x32 = x.astype(np.complex64).astype(np.complex128)
    sys_ = SH.ShearletSystem(grid)
    S = SO.shearlets_3d(grid)
    sys_.weights(_to_dev(x, dev))
    We_ref = SO.shearlet_weights(x32, S)
    # the weights come back through the regulariser
    z = crandn(rng, grid)
    z32 = z.astype(np.complex64).astype(np.complex128)
    Rg = sys_.regularize(_to_dev(z, dev), 0.7).cpu().numpy()
    Rr = SO.regularize_shearlet(z32, S, We_ref, 1.4, 1.0)  # (la p / 2) with la = 1.4, p = 1
    assert rel_l2(Rg, Rr) < 2e-5, rel_l2(Rg, Rr)
    e_r = SO.trace_samples_regularizer([z32.real], S, We_ref)[0]
    zr = _to_dev(z.real.astype(np.complex64), dev)
    e_g = sys_.energy(zr)
    assert abs(e_g / e_r - 1) < 2e-5, (e_g, e_r)
    Sm = crandn(rng, grid + (3,))
    W = rng.uniform(0.2, 1.0, grid[:2])
    Wd = torch.from_numpy(W.astype(np.float32).reshape(-1, order="F")).to(dev)
    f_g = sys_.fidelity(_to_dev(Sm, dev), Wd, zr)
    f_r = SO.trace_samples_fidelity([z32.real], Sm.astype(np.complex64).astype(np.complex128), W)[0]
    assert abs(f_g / f_r - 1) < 2e-5, (f_g, f_r)
    sys_.close()


@pytest.mark.gpu
@pytest.mark.parametrize("NY", [None, (32, 28)])
def test_cssolver_matches_oracle(dev, NY):
    """CSsolver on the GPU against the oracle; NY: data acquired on a smaller
    phase-encode grid (SENSE folding in the operators and the fidelity trace)."""
    from disorder_amd import ops
    from disorder_amd import sampling as SA
    from oracle import disorder_oracle as O
    from tests.helpers import checkered_sampling, smooth_volume
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(4)
    N, S, C = N3, 16, 4
    G = N[:2] if NY is None else NY
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 7)
    H = checkered_sampling(rng, G[0], G[1], S, 4, 4)
    k1, k2, off = SA.harmonic_sampling(H, S)
    T = np.zeros((S, 6))
    T[:, :3] = rng.uniform(-0.5, 0.5, (S, 3))
    We = rng.uniform(0.5, 1.0, S)[np.repeat(np.arange(S), np.diff(off))]
    samp = O.Sampling(k1, k2, off, G[0], G[1], We)
    y = O.encode(x_gt, Sm, samp, T, N)
    y = y + 0.05 * np.sqrt(np.mean(np.abs(y) ** 2)) * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))
    y = y.astype(np.complex64).astype(np.complex128)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    x0, _, _ = O.cg_solver(y, Sm32, samp, T, N, nIt=6, tolType="None", tol=0)
    x0 = x0.astype(np.complex64).astype(np.complex128)
    Wg = SH.grid_weights(k1, k2, We, G).astype(np.float64)
    x_ref, la_ref = SO.cs_solver(y, Sm32, samp, T, N, x0, Wg, seed=9)
    E = ops.Encoding(_to_dev(Sm, dev), k1, k2, off, We=We, NY=NY)
    xg, la_g = SH.CSsolver(_to_dev(y, dev), E, T, _to_dev(x0, dev), _to_dev(Sm, dev), k1, k2, We, seed=9)
    E.close()
    assert len(la_g) == len(la_ref) == 2
    assert max(abs(a / b - 1) for a, b in zip(la_g, la_ref)) < 1e-3, (la_g, la_ref)
    assert rel_l2(xg.cpu().numpy(), x_ref) < 1e-4, rel_l2(xg.cpu().numpy(), x_ref)
