"""Robust per-shot weights (Reconstruction/solveXT.m:450-481).

CPU: the oracle's MATLAB prctile against hand-worked values of its
definition (sorted value i at 100 (i - 0.5) / n, linear in between, clamped);
the oracle's weights against the normal model they assume (one line per
shot, log residuals at exact normal quantiles: the robust mean and deviation
recover mu and sigma, so a shot planted t deviations out gets
(1 - Phi(t)) NSweeps / 0.05); invariances (a common residual scale, shot
relabelling); the product's host weights (numpy Hazen percentiles) against
the oracle's loops; and that in the oracle driver the weights find a
corrupted shot and lower the image error.
GPU: the driver with robust weights against the oracle driver on a scan with
a corrupted shot: same weights, image and motion."""
import math

import numpy as np
import pytest
from scipy.special import ndtr, ndtri

from disorder_amd import robust as RB
from disorder_amd import sampling as SA
from oracle import disorder_oracle as O
from tests.helpers import checkered_sampling, rel_l2


def test_prctile_hand_values():
    x = np.array([4.0, 1.0, 3.0, 2.0])  # sorted 1 2 3 4 at 12.5, 37.5, 62.5, 87.5 %
    got = O.prctile(x, [0, 12.5, 25, 50, 75, 87.5, 90, 100])
    assert np.allclose(got, [1, 1, 1.5, 2.5, 3.5, 4, 4, 4], atol=0, rtol=0)
    y = np.arange(1.0, 6.0)  # 1..5 at 10, 30, 50, 70, 90 %
    assert np.allclose(O.prctile(y, [10, 20, 35, 50, 85, 95]), [1, 1.5, 2.25, 3, 4.75, 5], atol=1e-15)
    assert O.prctile([7.0], [0, 50, 100]).tolist() == [7.0, 7.0, 7.0]


@pytest.mark.parametrize("p", [0.0, 3.0, 12.5, 18.75, 25.0, 50.0, 61.0, 99.0, 100.0])
def test_product_prctile_matches_oracle(p):
    rng = np.random.default_rng(int(p * 4))
    for n in (1, 2, 5, 17, 64):
        x = rng.normal(size=n)
        assert abs(RB.prctile(x, p) - O.prctile(x, [p])[0]) < 1e-13


def _normal_shots(S, mu=-3.0, sigma=0.4):
    return mu + sigma * ndtri((np.arange(S) + 0.5) / S)


@pytest.mark.parametrize("t", [3.9, 4.2, 4.6])
def test_weights_recover_normal_model(t):
    """One line per shot: every percentile of a shot is its one value, so
    z is (log r - mu_hat) / sigma_hat.  With the shots' log residuals at the
    normal quantiles of N(mu, sigma^2), mu_hat and sigma_hat match mu and
    sigma up to the discretisation of the 12.5/18.75/25 % percentiles."""
    S, mu, sigma = 4000, -3.0, 0.4
    v = _normal_shots(S, mu, sigma)
    v[1234] = mu + t * sigma
    We = O.robust_shot_weights(np.exp(v), np.arange(S), S, finer=False)
    expect = min((1 - ndtr(t)) * S / 0.05, 1.0)
    assert abs(We[1234] - expect) < 0.03 * expect + 1e-6, (We[1234], expect)
    # lines below the centre never lose weight; the bulk keeps weight 1
    assert (We[v < mu + 2 * sigma] == 1.0).all()


def test_weights_invariances():
    rng = np.random.default_rng(4)
    S, per = 24, 30
    st = np.repeat(np.arange(S), per)
    r = np.exp(rng.normal(0, 0.5, S * per))
    r[st == 7] *= 8.0
    for finer in (False, True):
        We = O.robust_shot_weights(r, st, S, finer=finer)
        assert We[7] < 1e-3 and (np.delete(We, 7) > 0.5).all()
        # a common scale of the residuals is a shift of their logs
        assert np.allclose(O.robust_shot_weights(r * 37.0, st, S, finer=finer), We, atol=1e-12)
        # relabelling the shots permutes the weights
        perm = rng.permutation(S)
        assert np.allclose(O.robust_shot_weights(r, perm[st], S, finer=finer)[perm], We, atol=1e-12)
        # a shot whose residuals sit below the others keeps weight 1
        r2 = r.copy()
        r2[st == 3] /= 8.0
        assert O.robust_shot_weights(r2, st, S, finer=finer)[3] == 1.0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_product_weights_match_oracle(seed):
    rng = np.random.default_rng(seed)
    S = int(rng.integers(8, 64))
    counts = rng.integers(1, 40, S)
    counts[rng.integers(0, S)] = 0  # a shot without lines
    st = np.repeat(np.arange(S), counts)
    st = st[rng.permutation(len(st))]
    r = np.exp(rng.normal(0, 0.7, len(st)))
    r[st == st[0]] *= 5.0
    for finer in (False, True):
        ref = O.robust_shot_weights(r, st, S, finer=finer)
        got, z = RB.shot_weights(r, st, S, finer=finer)
        assert np.abs(got - ref).max() < 1e-12
        assert np.isnan(z[counts == 0]).all() and (got[counts == 0] == 1.0).all()
    with pytest.raises(ValueError):
        RB.shot_weights(r, st, 0)
    with pytest.raises(ValueError):
        RB.shot_weights(r[:-1], st, S)
    with pytest.raises(ValueError):
        RB.shot_weights(r, st + S, S)


def _corrupted_scan(seed=31, N=(32, 24, 16), S=16, C=4, bad=5, gain=3.0):
    from tests.helpers import smooth_volume
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(seed)
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5)
    H = checkered_sampling(rng, N[0], N[1], S, 4, 4)
    k1, k2, off = SA.harmonic_sampling(H, S)
    T_gt = np.zeros((S, 6))
    T_gt[:, :3] = rng.uniform(-1, 1, (S, 3))
    T_gt[:, 3:] = rng.uniform(-2, 2, (S, 3)) * np.pi / 180
    samp = O.Sampling(k1, k2, off, N[0], N[1])
    y = O.encode(x_gt, Sm, samp, T_gt, N)
    sl = slice(off[bad], off[bad + 1])
    scale = np.sqrt(np.mean(np.abs(y) ** 2))
    y[sl] += gain * scale * (rng.normal(size=y[sl].shape) + 1j * rng.normal(size=y[sl].shape)) / math.sqrt(2)
    y += 0.01 * scale * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape)) / math.sqrt(2)
    y = y.astype(np.complex64).astype(np.complex128)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    return x_gt, Sm, Sm32, y, k1, k2, off, N, bad


def test_oracle_driver_rejects_corrupted_shot():
    x_gt, _, Sm32, y, k1, k2, off, N, bad = _corrupted_scan()
    info = {}
    # gibbs = 0: the Tukey window zeroes the k-space corners of the data (not of
    # the model), which raises the residuals of shots sampling the corners of
    # this full rectangular grid; a scanner's elliptical shutter never samples them
    xr, _ = O.pyramid_solve(y, Sm32, k1, k2, off, N, nlevels=2, n_outer=3, n_cg=4, robust=True, info=info, gibbs=0)
    xp, _ = O.pyramid_solve(y, Sm32, k1, k2, off, N, nlevels=2, n_outer=3, n_cg=4, robust=False, gibbs=0)
    We = info["We"]
    assert We[bad] < 0.05 and (np.delete(We, bad) > 0.5).all(), We
    assert rel_l2(xr, x_gt) < 0.8 * rel_l2(xp, x_gt), (rel_l2(xr, x_gt), rel_l2(xp, x_gt))


@pytest.mark.gpu
def test_driver_robust_weights_match_oracle(dev):
    from disorder_amd import pyramid as PY
    from tests.test_pyramid import _to_dev
    x_gt, Sm, Sm32, y, k1, k2, off, N, bad = _corrupted_scan()
    info_o = {}
    x_ref, T_ref = O.pyramid_solve(y, Sm32, k1, k2, off, N, nlevels=2, n_outer=3, n_cg=4, robust=True, info=info_o,
                                   gibbs=0)
    xg, Tg, info = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, nlevels=2, n_outer=3, n_cg=4,
                              robust=True, gibbs=0)
    assert np.abs(info["We"] - info_o["We"]).max() < 1e-3, (info["We"], info_o["We"])
    assert info["We"][bad] < 0.05
    assert rel_l2(xg.cpu().numpy(), x_ref) < 1e-4
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01
