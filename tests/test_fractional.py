"""Fractional finite-difference motion estimation (E.Fd): the weights of
Methods/buildFilter.m 'FractionalFiniteDiscreteIso' built per level by
Reconstruction/solveXT.m:353-354 and applied by Reconstruction/LMsolver.m
(:88-91 residual, :101-104 Jacobians, :186-189 trial residuals) after the
unnormalised readout DFT of the ROI.

CPU: the oracle's filter against closed forms (order 0 is all ones; on an
axis the filter is |1 - exp(-2 pi i m / (sp n))|^order, the transfer
function of a forward difference at spacing sp, so the FFT-order layout, the
2 pi / (sp N) scale and the half-angle sine are fixed; corners beyond pi are
clipped to 2^order; the zero frequency takes the largest off-zero value;
powers multiply); the product's per-sample weights against the oracle's
literal grid; the oracle's weighted LM sums against Parseval (unit weights
give numel(ROI) times the plain sums) and against central differences of
the weighted energy (the Jacobians are weighted like the residual).
GPU: dis_lm_partials and one dis_lm call with the weights against the
oracle (whole readout and the 'H' ROI), and the pyramid driver with
fraction_order = 1 against the oracle driver."""
import math

import numpy as np
import pytest

from disorder_amd import pyramid as PY
from disorder_amd import sampling as SA
from oracle import disorder_oracle as O
from tests.helpers import checkered_sampling, crandn, random_motion, rel_l2, smooth_volume


def _axis_transfer(n, sp, order):
    m = np.where(np.arange(n) < (n + 1) // 2, np.arange(n), np.arange(n) - n)
    with np.errstate(divide="ignore"):
        return np.abs(1.0 - np.exp(-2j * np.pi * m / (sp * n))) ** order


@pytest.mark.parametrize("n,sp,order", [(8, 1.0, 1.0), (9, 1.0, 1.0), (12, 2.0, 0.5), (7, 4.0, 2.0), (10, 1.0, -0.5)])
def test_filter_on_an_axis_is_the_forward_difference_transfer(n, sp, order):
    for d in range(3):
        N = [1, 1, 1]
        N[d] = n
        F = O.fractional_filter(N, [sp] * 3, order).reshape(-1)
        ref = _axis_transfer(n, sp, order)
        assert np.allclose(F[1:], ref[1:], rtol=1e-12, atol=0), (d, F, ref)
        assert F[0] == pytest.approx(np.max(ref[1:]), rel=1e-13)  # zero frequency: the largest of the rest


def test_filter_order_zero_clip_and_powers():
    assert np.array_equal(O.fractional_filter([6, 5, 4], [1, 1, 1], 0.0), np.ones((6, 5, 4)))
    # the corner (Nyquist on every axis) lies at radius pi sqrt(3): clipped to 2
    F1 = O.fractional_filter([8, 8, 8], [1, 1, 1], 1.0)
    assert F1[4, 4, 4] == pytest.approx(2.0, abs=1e-15)
    assert F1[0, 0, 0] == pytest.approx(2.0, abs=1e-15)
    # radial: (1, 1, 0) and (0, 1, 1) have the same radius
    assert F1[1, 1, 0] == pytest.approx(F1[0, 1, 1], rel=1e-14)
    assert F1[1, 1, 0] == pytest.approx(2 * math.sin(math.sqrt(2) * 2 * math.pi / 8 / 2), rel=1e-14)
    Fa = O.fractional_filter([6, 5, 7], [2, 1, 1.5], 0.7)
    Fb = O.fractional_filter([6, 5, 7], [2, 1, 1.5], 0.6)
    Fab = O.fractional_filter([6, 5, 7], [2, 1, 1.5], 1.3)
    assert np.allclose((Fa * Fb).reshape(-1)[1:], Fab.reshape(-1)[1:], rtol=1e-13)


@pytest.mark.parametrize("n,nF,sp,order", [((16, 12, 10), 7, (2, 2, 2), 1.0), ((17, 9, 8), 8, (1, 1, 1), 0.5),
                                           ((8, 6, 5), 5, (4, 4, 3.2), -0.5), ((16, 12, 10), 10, (1, 1, 1), 2.0),
                                           ((9, 1, 6), 6, (1, 1, 1), 1.0), ((16, 12, 10), 10, (1, 1, 1), 0.0)])
def test_product_weights_match_oracle_grid(n, nF, sp, order):
    rng = np.random.default_rng(sum(n) + nF)
    k1 = rng.integers(0, n[0], 200)
    k2 = rng.integers(0, n[1], 200)
    k1[0] = k2[0] = 0
    ref = O.fractional_filter([n[0], n[1], nF], sp, order)[k1, k2, :]
    got = PY.fractional_weights(n, nF, sp, order, k1, k2)
    assert got.dtype == np.float32 and got.shape == (200, nF)
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 2e-7


def _small_lm_problem(rng, n=(12, 10, 8), C=2, S=4):
    x = smooth_volume(rng, n) + 0.1 * crandn(rng, n)
    Sm = crandn(rng, n + (C,)) * 0.5
    H = checkered_sampling(rng, n[0], n[1], S, 2, 2)
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off[:S + 1], n[0], n[1])
    Tgt = random_motion(rng, S, 3, 1)
    y = O.encode(x, Sm, samp, Tgt, n) + 0.01 * crandn(rng, (samp.Ns, n[2], C))
    return x, Sm, samp, y, Tgt + random_motion(rng, S, 1.0, 0.3)


@pytest.mark.parametrize("nF", [None, (3, 8)])
def test_oracle_weighted_lm_parseval_and_gradient(nF):
    rng = np.random.default_rng(5)
    x, Sm, samp, y, T0 = _small_lm_problem(rng)
    n = x.shape
    L = n[2] if nF is None else nF[1] - nF[0]
    ones = np.ones((samp.Ns, L))
    E0, G0, H0 = O.lm_partials(y, x, Sm, samp, T0, n, nF=nF)
    E1, G1, H1 = O.lm_partials(y, x, Sm, samp, T0, n, nF=nF, Fd=ones)
    assert np.allclose(E1, L * E0, rtol=1e-12) and np.allclose(G1, L * G0, rtol=1e-10)
    assert np.allclose(H1, L * H0, rtol=1e-10)
    # order 1 weights: the gradient is half the derivative of the energy
    F3 = O.fractional_filter([n[0], n[1], L], [1, 1, 1], 1.0)
    Fd = O.fractional_samples(F3, samp)
    E, G, _ = O.lm_partials(y, x, Sm, samp, T0, n, nF=nF, Fd=Fd)
    NElY = y.size
    h = 1e-5
    for g in range(6):
        Tp, Tm = T0.copy(), T0.copy()
        Tp[:, g] += h
        Tm[:, g] -= h
        ep = O.lm_state_energy(y, x, Sm, samp, Tp, n, range(4), nF=nF, Fd=Fd)
        em = O.lm_state_energy(y, x, Sm, samp, Tm, n, range(4), nF=nF, Fd=Fd)
        fd = np.array([(ep[s] - em[s]) / (2 * h) for s in range(4)])
        assert np.allclose(fd, 2 * G[:, g] / NElY, rtol=1e-5, atol=1e-12), (g, fd, 2 * G[:, g] / NElY)
    # the weights change the sums (not a no-op)
    assert rel_l2(E, L * E0) > 0.05


def test_weighted_trial_energy_equals_partials_energy():
    rng = np.random.default_rng(6)
    x, Sm, samp, y, T0 = _small_lm_problem(rng)
    n = x.shape
    Fd = O.fractional_samples(O.fractional_filter([n[0], n[1], n[2]], [1, 1, 1], 0.5), samp)
    E, _, _ = O.lm_partials(y, x, Sm, samp, T0, n, Fd=Fd)
    Et = O.lm_state_energy(y, x, Sm, samp, T0, n, range(4), Fd=Fd)
    assert np.allclose(E, [Et[s] for s in range(4)], rtol=1e-12)


def test_level_rejects_weights_on_folded_data():
    with pytest.raises(ValueError):
        O.pyramid_solve(np.zeros((4, 4, 1)), np.zeros((4, 4, 4, 1)), [0], [0], [0, 1], (4, 4, 4), nlevels=1,
                        NY=(2, 4), fraction_order=1.0)


# ---------------------------------------------------------------- GPU
def _to_dev(a, dev):
    import torch
    import disorder_amd as D
    return D.colmajor(torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64))).to(dev))


@pytest.mark.gpu
@pytest.mark.parametrize("n,C,S", [((16, 12, 8), 3, 4), ((32, 24, 10), 4, 4)])
@pytest.mark.parametrize("fat,order", [(None, 1.0), ("H", 1.0), ("H", 0.5)])
def test_weighted_lm_partials_and_step_match_oracle(dev, n, C, S, fat, order):
    """dis_lm_partials and dis_lm with E.Fd set against the oracle's weighted
    LMsolver (energies 1e-5, gradients and Hessians 1e-4, motion 1e-5)."""
    import disorder_amd as D
    rng = np.random.default_rng(43)
    x = smooth_volume(rng, n) + 0.1 * crandn(rng, n)
    Sm = crandn(rng, n + (C,)) * 0.5
    H = checkered_sampling(rng, n[0], n[1], S, 2, 2)
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off[:S + 1], n[0], n[1])
    Tgt = random_motion(rng, S, 3, 1)
    y = O.encode(x, Sm, samp, Tgt, n) + 0.01 * crandn(rng, (samp.Ns, n[2], C))
    T0 = Tgt + random_motion(rng, S, 1.0, 0.3)
    nF = None if fat is None else O.readout_roi(n[2], fat)
    L = n[2] if nF is None else nF[1] - nF[0]
    Fd = O.fractional_samples(O.fractional_filter([n[0], n[1], L], [1, 1, 1], order), samp)
    E = D.Encoding(_to_dev(Sm, dev), samp.k1, samp.k2, samp.off, Nfull=n)
    E.set_readout_roi(nF)
    E.set_lm_filter(PY.fractional_weights(n, L, (1, 1, 1), order, samp.k1, samp.k2))
    xd, yd = _to_dev(x, dev), _to_dev(y, dev)
    En, dG, dH = D.lm_partials(yd, E, xd, T0)
    Er, dGr, dHr = O.lm_partials(y, x, Sm, samp, T0, n, nF=nF, Fd=Fd)
    assert rel_l2(En, Er) < 1e-5
    assert rel_l2(dG, dGr) < 1e-4
    assert rel_l2(dH, dHr) < 1e-4
    w = np.full(S, 1e-3)
    cT = np.zeros(S, bool)
    Tg, wg, cg, Eg, _ = D.LMsolver(yd, E, xd, T0, w, cT, winit=1e-3, tL=(0.05, 0.02))
    Tr, wr, cr, Ern = O.lm_solver(y, x, Sm, samp, T0, n, w, cT, 1e-3, (0.05, 0.02), nF=nF, Fd=Fd)
    assert np.allclose(wg, wr) and np.array_equal(cg, cr)
    assert np.abs(Tg - Tr).max() < 1e-5
    assert np.allclose(Eg, Ern, rtol=1e-5, atol=1e-12)
    # the weights must match the ROI length
    E.set_readout_roi(None if nF is not None else (0, n[2] - 1))
    from disorder_amd._lib import DisorderError
    with pytest.raises(DisorderError, match="readout ROI"):
        D.lm_partials(yd, E, xd, T0)
    E.close()


@pytest.mark.gpu
def test_pyramid_driver_fractional_matches_oracle(dev):
    """Two levels, two outer iterations, fraction_order = 1 (E.Fd on every
    LMsolver call, readout ROI 'H'): the GPU driver and the oracle driver
    reach the same image (NRMSE <= 1e-4) and motion (0.01 deg, 0.01 voxel)."""
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(24)
    N = (32, 24, 16)
    S, C = 16, 4
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5)
    H = checkered_sampling(rng, N[0], N[1], S, 4, 4)
    k1, k2, off = SA.harmonic_sampling(H, S)
    T_gt = np.zeros((S, 6))
    T_gt[:, :3] = rng.uniform(-1, 1, (S, 3))
    T_gt[:, 3:] = rng.uniform(-2, 2, (S, 3)) * np.pi / 180
    samp = O.Sampling(k1, k2, off, N[0], N[1])
    y = O.encode(x_gt, Sm, samp, T_gt, N).astype(np.complex64).astype(np.complex128)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    kw = dict(nlevels=2, n_outer=2, n_cg=4, fraction_order=1.0)
    x_ref, T_ref = O.pyramid_solve(y, Sm32, k1, k2, off, N, **kw)
    xg, Tg, info = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, **kw)
    xg = xg.cpu().numpy()
    assert rel_l2(xg, x_ref) < 1e-4, rel_l2(xg, x_ref)
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01
    # the weights change the estimate (not a no-op)
    _, T_plain = O.pyramid_solve(y, Sm32, k1, k2, off, N, nlevels=2, n_outer=2, n_cg=4)
    assert np.abs(T_plain - T_ref).max() > 1e-4
