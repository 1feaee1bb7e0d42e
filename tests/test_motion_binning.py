"""Motion binning (Methods/compressMotion.m:27-80).

CPU: the oracle's Haar dwt against MATLAB's documented dwt([1 2 3 4],'haar')
example; wavedec against the orthonormal Haar basis built from its
definition (scaled box and step functions, MATLAB's coefficient order
[a_NL, d_NL, ..., d_1]) for power-of-two lengths; perfect reconstruction
for every length 1..70 (odd lengths take the 'sym' boundary); binning of
constant, dyadic step and spiked motion to the bins the thresholds imply;
the product's binning (vectorised numpy) against the oracle's loops.
GPU: an encoding over bins with the bin motion equals the per-shot
encoding with each shot given its bin's motion (E^H E and E x)."""
import math

import numpy as np
import pytest

from disorder_amd import motion as MB
from oracle import disorder_oracle as O
from tests.helpers import rel_l2


def test_dwt_matlab_example():
    a, d = O.haar_dwt([1.0, 2.0, 3.0, 4.0])
    assert np.allclose(a, [2.1213, 4.9497], atol=1e-4) and np.allclose(d, [-0.7071, -0.7071], atol=1e-4)
    a, d = O.haar_dwt([1.0, 2.0, 3.0])  # 'sym' repeats the last sample
    assert np.allclose(a, [3 / math.sqrt(2), 6 / math.sqrt(2)]) and np.allclose(d, [-1 / math.sqrt(2), 0.0])


def _haar_basis(J):
    """Rows: the constant 2^(-J/2), then for j = J..1 the steps of width 2^j
    (+1 on the first half, -1 on the second) scaled by 2^(-j/2)."""
    n = 2 ** J
    rows = [np.full(n, 2.0 ** (-J / 2))]
    for j in range(J, 0, -1):
        w = 2 ** j
        for k in range(n // w):
            r = np.zeros(n)
            r[k * w:k * w + w // 2] = 1.0
            r[k * w + w // 2:(k + 1) * w] = -1.0
            rows.append(r * 2.0 ** (-j / 2))
    return np.array(rows)


@pytest.mark.parametrize("J", [1, 2, 3, 5, 6])
def test_wavedec_is_orthonormal_haar(J):
    H = _haar_basis(J)
    assert np.allclose(H @ H.T, np.eye(2 ** J), atol=1e-12)
    x = np.random.default_rng(J).normal(size=2 ** J)
    C, L = O.wavedec_haar(x, J)
    assert np.allclose(C, H @ x, atol=1e-12)
    assert L == [1] + [2 ** (j - 1) for j in range(1, J + 1)] + [2 ** J]


def test_waverec_perfect_reconstruction():
    rng = np.random.default_rng(7)
    for n in range(1, 71):
        x = rng.normal(size=n)
        NL = int(math.floor(math.log2(n))) if n > 1 else 0
        C, L = O.wavedec_haar(x, NL)
        assert np.abs(O.waverec_haar(C, L) - x).max() < 1e-12, n
        assert np.abs(MB.haar_denoise(x[:, None], [0.0], NL)[:, 0] - x).max() < 1e-12, n


def _off(S, rng):
    return np.concatenate([[0], np.cumsum(rng.integers(1, 9, S))])


def test_binning_of_known_motion():
    rng = np.random.default_rng(3)
    res = math.sqrt(2.0)
    th_t, th_r = 0.05 * res, 0.02 * res * math.pi / 180
    S = 32
    off = _off(S, rng)
    # constant motion above the limits: one bin with that motion
    T = np.tile([0.8, -0.5, 0.3, 0.01, -0.02, 0.015], (S, 1))
    Tb, offb, bos = O.compress_motion(T, off, res)
    assert len(Tb) == 1 and np.allclose(Tb[0], T[0], atol=1e-12) and offb.tolist() == [0, off[-1]]
    # four dyadic blocks of 8 shots, jitter well below the limits: 4 bins at the block means
    levels = rng.normal(0, 2, (4, 6))
    levels[:, 3:] *= math.pi / 180
    jit = rng.uniform(-1, 1, (S, 6)) * np.r_[[th_t] * 3, [th_r] * 3] * 0.02
    T = np.repeat(levels, 8, axis=0) + jit
    Tb, offb, bos = O.compress_motion(T, off, res)
    assert bos.tolist() == np.repeat(np.arange(4), 8).tolist()
    assert offb.tolist() == off[::8].tolist()
    assert np.allclose(Tb, T.reshape(4, 8, 6).mean(axis=1), atol=1e-12)
    # a spike on one shot far above the limits keeps that shot in a bin of its own
    T2 = np.repeat(levels, 8, axis=0)
    T2[13, 0] += 5.0
    Tb, offb, bos = O.compress_motion(T2, off, res)
    assert (bos == bos[13]).sum() == 1 and abs(Tb[bos[13], 0] - T2[13, 0]) < 1e-9
    # one shot: unchanged
    Tb, offb, bos = O.compress_motion(T[:1], off[:2], res)
    assert np.array_equal(Tb, T[:1]) and offb.tolist() == off[:2].tolist()
    # the approximation coefficient is thresholded too: 4 shots at a common
    # motion c have a = 2c; below the limit the bin's motion is zero
    c = np.r_[[0.3 * th_t] * 3, [0.3 * th_r] * 3]
    Tb, offb, bos = O.compress_motion(np.tile(c, (4, 1)), off[:5], res)
    assert len(Tb) == 1 and np.array_equal(Tb[0], np.zeros(6))
    Tb, _, _ = O.compress_motion(np.tile(2 * c, (4, 1)), off[:5], res)
    assert np.allclose(Tb[0], 2 * c, atol=1e-15)
    # without thresholds a run is closed by the first shot more than 1e-6 from
    # the run's first shot (not from its neighbour): a ramp of 0.6e-6 steps
    R = np.zeros((8, 6))
    R[:, 0] = 0.6e-6 * np.arange(8)
    _, _, bos = O.compress_motion(R, off[:9], res, traLimX=(0.0, 0.0))
    assert bos.tolist() == [0, 0, 1, 1, 2, 2, 3, 3]
    assert MB.bin_motion(R, off[:9], res, traLimX=(0.0, 0.0))[2].tolist() == bos.tolist()


@pytest.mark.parametrize("seed", range(4))
def test_product_binning_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    for S in (2, 3, 7, 16, 45, 64, 100):
        T = np.repeat(rng.normal(0, 1, (S // 3 + 1, 6)), 3, axis=0)[:S]
        T[:, 3:] *= math.pi / 180
        T += rng.normal(0, 0.02, (S, 6)) * np.r_[[1.0] * 3, [0.3 * math.pi / 180] * 3]
        off = _off(S, rng)
        a = O.compress_motion(T, off, math.sqrt(2.0))
        b = MB.bin_motion(T, off, math.sqrt(2.0))
        assert len(a[0]) == len(b[0]), (S, len(a[0]), len(b[0]))
        assert np.abs(a[0] - b[0]).max() < 1e-12
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    with pytest.raises(ValueError):
        MB.bin_motion(np.zeros((4, 6)), np.arange(4), 1.0)


@pytest.mark.gpu
def test_binned_encoding_equals_per_shot(dev):
    from disorder_amd import ops
    from disorder_amd import sampling as SA
    from tests.helpers import checkered_sampling, crandn
    from tests.test_pyramid import _to_dev
    rng = np.random.default_rng(12)
    N, S, C = (32, 24, 20), 16, 4
    H = checkered_sampling(rng, N[0], N[1], S, 4, 4)
    k1, k2, off = SA.harmonic_sampling(H, S)
    levels = rng.normal(0, 1.5, (4, 6))
    levels[:, 3:] *= math.pi / 180
    T = np.repeat(levels, 4, axis=0) + rng.normal(0, 1e-3, (S, 6)) * np.r_[[1.0] * 3, [0.01] * 3]
    Tb, offb, bos = MB.bin_motion(T, off, math.sqrt(2.0))
    assert len(Tb) == 4, Tb
    Sm = _to_dev(crandn(rng, N + (C,)), dev)
    x = _to_dev(crandn(rng, N), dev)
    We = rng.uniform(0.2, 1.0, int(off[-1]))
    Es = ops.Encoding(Sm, k1, k2, off, Nfull=N, We=We)
    Eb = ops.Encoding(Sm, k1, k2, offb, Nfull=N, We=We)
    ys = ops.encode(x, Es, Tb[bos]).cpu().numpy()
    yb = ops.encode(x, Eb, Tb).cpu().numpy()
    assert rel_l2(yb, ys) < 1e-5
    hs = ops.encodedecode(x, Es, Tb[bos]).cpu().numpy()
    hb = ops.encodedecode(x, Eb, Tb).cpu().numpy()
    assert rel_l2(hb, hs) < 1e-5
    Es.close()
    Eb.close()
