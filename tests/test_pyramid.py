"""Pyramid levels (Methods/resampling.m, Methods/downsampleOperators.m).

CPU: the oracle's Fourier resampling against closed forms (band-limited
complex exponentials resample exactly; up then down is the identity), the
k-space mask crop, and the host level sampling (bit-exact against cropping the
mask and re-running the harmonic sampling, as the reference does).
GPU: dis_resample against the oracle (rel-L2 <= 1e-5) for down/up/odd sizes
and batches, the per-level compact data against the oracle's literal
image-space route, and the preconditioner."""
import math

import numpy as np
import pytest

from disorder_amd import sampling as SA
from oracle import disorder_oracle as O
from tests.helpers import checkered_sampling, crandn, rel_l2


@pytest.mark.parametrize("N,n", [(16, 8), (16, 9), (15, 8), (15, 6), (12, 12), (8, 16), (7, 12)])
def test_resample_band_limited_exponentials_exact(N, n):
    nmin = min(N, n)
    j = np.arange(N)
    jn = np.arange(n)
    for f in range(-(nmin // 2), (nmin - 1) // 2 + 1):
        if N % 2 == 0 and f == -(N // 2) and n > N:
            continue  # the input's Nyquist bin is not a resolvable exponential
        x = np.exp(2j * np.pi * f * j / N)[:, None, None]
        y = O.resample_fourier(x, (n, 1, 1))
        ref = np.exp(2j * np.pi * f * jn / n)[:, None, None]
        assert np.abs(y - ref).max() < 1e-12, (N, n, f)
    # frequencies outside the kept band are removed
    fo = (nmin - 1) // 2 + 1
    if fo < (N + 1) // 2 and N > n:
        x = np.exp(2j * np.pi * fo * j / N)[:, None, None]
        assert np.abs(O.resample_fourier(x, (n, 1, 1))).max() < 1e-12


def test_resample_kspace_mask_crop():
    H = np.arange(1, 8 * 6 + 1).reshape(8, 6, order="F")
    Hc = O.resample_kspace(H, (4, 3))
    # fftshift, centre crop, ifftshift of the DC-first plane keeps the
    # frequencies [-2, 1] x [-1, 1]
    f1 = [0, 1, -2, -1]
    f2 = [0, 1, -1]
    for a, u in enumerate(f1):
        for b, v in enumerate(f2):
            assert Hc[a, b] == H[u % 8, v % 6]


@pytest.mark.parametrize("N12,n12,t", [((16, 12), (8, 6), (4, 3)), ((16, 12), (4, 3), (4, 3)),
                                       ((15, 9), (8, 5), (3, 3)), ((32, 24), (16, 12), (4, 4))])
def test_level_sampling_bit_exact(N12, n12, t):
    rng = np.random.default_rng(sum(N12))
    S = t[0] * t[1]
    H = checkered_sampling(rng, N12[0] - N12[0] % t[0], N12[1] - N12[1] % t[1], S, *t)
    Hf = np.zeros(N12, np.int64)
    Hf[:H.shape[0], :H.shape[1]] = H
    k1, k2, off = SA.harmonic_sampling(Hf, S)
    rows, l1, l2, loff = SA.level_sampling(k1, k2, off, N12, n12)
    Hl = O.resample_kspace(Hf, n12)
    r1, r2, roff, _ = O.build_harmonic_sampling(Hl, S)
    ns = roff[S]
    assert np.array_equal(l1, r1[:ns]) and np.array_equal(l2, r2[:ns])
    assert np.array_equal(loff, roff[:S + 1])
    # rows point at the same phase-encode positions at full resolution
    assert np.array_equal(SA.signed_freq(k1[rows], N12[0]) % n12[0], l1)


# ---------------------------------------------------------------- GPU
def _to_dev(a, dev):
    import disorder_amd as D
    import torch
    t = torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64)))
    return D.colmajor(t.to(dev))


@pytest.mark.gpu
@pytest.mark.parametrize("nin,nout,B", [((32, 24, 32), (16, 12, 16), 1), ((16, 12, 16), (32, 24, 32), 2),
                                        ((15, 14, 9), (8, 7, 5), 3), ((64, 48, 64), (32, 24, 32), 1),
                                        ((20, 18, 16), (20, 9, 16), 2), ((8, 6, 8), (8, 6, 8), 1),
                                        ((256, 8, 4), (128, 8, 4), 1),
                                        # pyramid of a 176 axis (176 -> 88 -> 44 -> 22 -> 11) and 13/26
                                        ((176, 26, 13), (88, 13, 7), 1), ((22, 13, 176), (11, 7, 88), 2),
                                        ((11, 13, 22), (22, 26, 44), 1)])
def test_resample_matches_oracle(dev, nin, nout, B):
    import disorder_amd as D
    rng = np.random.default_rng(sum(nin) + B)
    x = crandn(rng, nin + (B,))
    got = D.resampling(_to_dev(x, dev), nout).cpu().numpy()
    ref = O.resample_fourier(x, nout + (B,))
    assert rel_l2(got, ref) < 1e-5


@pytest.mark.gpu
def test_level_data_matches_oracle_route(dev):
    import disorder_amd as D
    rng = np.random.default_rng(3)
    N = (16, 12, 10)
    n = (8, 6, 5)
    S, C = 12, 3
    H = checkered_sampling(rng, N[0], N[1], S, 4, 3)
    k1, k2, off = SA.harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off, N[0], N[1])
    yc = crandn(rng, (len(k1), N[2], C))
    yl_ref, r1, r2, roff = O.level_data(yc, samp, H, N, n)
    rows, l1, l2, loff = SA.level_sampling(k1, k2, off, N[:2], n[:2])
    assert np.array_equal(l1, r1) and np.array_equal(l2, r2)
    scale = math.sqrt(n[0] * n[1] / (N[0] * N[1]))
    g = D.gather_samples(_to_dev(yc, dev), rows, scale)
    got = D.resampling(g, (len(rows), n[2], C))
    assert rel_l2(got.cpu().numpy(), yl_ref) < 1e-5


@pytest.mark.gpu
def test_coil_precond(dev):
    import disorder_amd as D
    rng = np.random.default_rng(4)
    Sm = crandn(rng, (9, 7, 5, 4))
    Se = D.coil_precond(_to_dev(Sm, dev)).cpu().numpy()
    ref = 1.0 / (np.sum(np.abs(Sm.astype(np.complex64)) ** 2, axis=-1) + 1e-9)
    assert np.abs(Se - ref).max() / ref.max() < 1e-6


def _coil_maps_np(n, C, seed):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.linspace(-1, 1, v) for v in n], indexing="ij")
    maps = []
    for c in range(C):
        a = 2 * np.pi * c / C
        ctr = (0.8 * np.cos(a), 0.8 * np.sin(a), rng.uniform(-0.3, 0.3))
        r2 = sum((gi - ci) ** 2 for gi, ci in zip(g, ctr))
        maps.append(np.exp(-r2 / 1.2) * np.exp(1j * (0.5 * g[0] * np.cos(a) + 0.5 * g[1] * np.sin(a))))
    Sm = np.stack(maps, -1)
    return Sm / np.sqrt(np.sum(np.abs(Sm) ** 2, -1, keepdims=True))


@pytest.mark.gpu
def test_pyramid_driver_matches_oracle(dev):
    """Two levels, two outer iterations each (CG then LM): the GPU driver and
    the FP64 oracle driver reach the same image (NRMSE <= 1e-4) and motion
    (0.01 deg, 0.01 voxel) from the same data."""
    import disorder_amd as D
    from disorder_amd import pyramid as PY
    from tests.helpers import smooth_volume
    rng = np.random.default_rng(21)
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
    x_ref, T_ref = O.pyramid_solve(y, Sm32, k1, k2, off, N, nlevels=2, n_outer=2, n_cg=4)
    xg, Tg, info = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, nlevels=2, n_outer=2, n_cg=4)
    xg = xg.cpu().numpy()
    assert rel_l2(xg, x_ref) < 1e-4, rel_l2(xg, x_ref)
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01
    assert [lv["n"] for lv in info["levels"]] == [[16, 12, 8], [32, 24, 16]]


@pytest.mark.gpu
def test_pyramid_driver_coil_split_matches_oracle(dev):
    """eivaS = (4, 2, 3) of 4 coils: image solves while estimating on the
    first 3 coils, LMsolver on the first 2, the closing CG on all 4; the GPU
    driver and the oracle driver agree (image NRMSE <= 1e-4, motion 0.01)."""
    from disorder_amd import pyramid as PY
    from tests.helpers import smooth_volume
    rng = np.random.default_rng(22)
    N = (32, 24, 16)
    S, C = 16, 4
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 6)
    H = checkered_sampling(rng, N[0], N[1], S, 4, 4)
    k1, k2, off = SA.harmonic_sampling(H, S)
    T_gt = np.zeros((S, 6))
    T_gt[:, :3] = rng.uniform(-1, 1, (S, 3))
    T_gt[:, 3:] = rng.uniform(-2, 2, (S, 3)) * np.pi / 180
    samp = O.Sampling(k1, k2, off, N[0], N[1])
    y = O.encode(x_gt, Sm, samp, T_gt, N).astype(np.complex64).astype(np.complex128)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    kw = dict(nlevels=2, n_outer=3, n_cg=4, final_cg=6, eivaS=(4, 2, 3))
    x_ref, T_ref = O.pyramid_solve(y, Sm32, k1, k2, off, N, **kw)
    xg, Tg, info = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, **kw)
    assert rel_l2(xg.cpu().numpy(), x_ref) < 1e-4, rel_l2(xg.cpu().numpy(), x_ref)
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01


def test_tukey_window_closed_form():
    """buildFilter 'tukeyIso' (DC first): 1 at DC and below pi (1 - g), the
    raised cosine between, 0 from pi on (the k-space corners)."""
    N, g = (16, 12, 10), 0.25
    H = O.tukey_iso(N, g)
    assert H[0, 0, 0] == 1.0
    # Nyquist along dim 1: |k| = pi -> 0; the corner (pi, pi, pi) -> 0
    assert H[8, 0, 0] == 0.0 and H[8, 6, 5] == 0.0
    # a point at radius r on axis 1: k = 2 pi f / 16
    for f in range(1, 8):
        r = 2 * np.pi * f / 16
        ref = 1.0 if r < np.pi * (1 - g) else 0.5 * (1 + np.cos(np.pi * (r - np.pi * (1 - g)) / (g * np.pi)))
        assert abs(H[f, 0, 0] - ref) < 1e-12 and abs(H[16 - f, 0, 0] - ref) < 1e-12
    # radial in radians: f = 4 of 16 and f = 3 of 12 are both at pi/2
    assert H[4, 0, 0] == H[0, 3, 0] and H[12, 0, 0] == H[0, 9, 0]
    # gibbs = 0: a hard spherical shutter at pi
    H0 = O.tukey_iso(N, 0.0)
    assert set(np.unique(H0)) <= {0.0, 1.0} and H0[7, 0, 0] == 1.0 and H0[8, 0, 0] == 0.0


def test_cubic_bspline_closed_form():
    """6 sinc(f/N)^4 / (4 + 2 cos(2 pi f/N)) per dimension, DC first: 1 at DC,
    separable, even in f, and 6 sinc(1/2)^4/2 = 48/pi^4 at the Nyquist bin."""
    N = (16, 12, 10)
    H = O.cubic_bspline(N)
    assert abs(H[0, 0, 0] - 1) < 1e-15
    assert abs(H[8, 0, 0] - 48 / np.pi ** 4) < 1e-12
    assert abs(H[3, 5, 2] - H[3, 0, 0] * H[0, 5, 0] * H[0, 0, 2]) < 1e-15
    assert abs(H[3, 0, 0] - H[13, 0, 0]) < 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize("n,C,g", [((32, 24, 16), 3, 0.05), ((20, 18, 15), 2, 0.05), ((64, 48, 40), 4, 0.05),
                                   ((32, 24, 16), 2, 2.0)])
def test_filter_rows_matches_oracle(dev, n, C, g):
    from disorder_amd import ops
    rng = np.random.default_rng(30)
    Ns = 300
    l1 = rng.integers(0, n[0], Ns)
    l2 = rng.integers(0, n[1], Ns)
    y = crandn(rng, (Ns, n[2], C))
    ref = O.filter_level_rows(y.astype(np.complex64).astype(np.complex128), l1, l2, n, g)
    yd = _to_dev(y, dev)
    ops.filter_rows(yd, l1, l2, n, g)
    assert rel_l2(yd.cpu().numpy(), ref) < 1e-5


@pytest.mark.gpu
def test_pyramid_driver_accel_matches_oracle(dev):
    """accel = 1 (pyramidPlan.m:23 default): the full-resolution level runs
    only its closing CG; with the Gibbs window on (default 0.05) the GPU
    driver and the oracle driver agree (image 1e-4, motion 0.01)."""
    from disorder_amd import pyramid as PY
    from tests.helpers import smooth_volume
    rng = np.random.default_rng(23)
    N = (32, 24, 16)
    S, C = 16, 4
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5)
    H = checkered_sampling(rng, N[0], N[1], S, 4, 4)
    k1, k2, off = SA.harmonic_sampling(H, S)
    T_gt = np.zeros((S, 6))
    T_gt[:, :3] = rng.uniform(-1, 1, (S, 3))
    samp = O.Sampling(k1, k2, off, N[0], N[1])
    y = O.encode(x_gt, Sm, samp, T_gt, N).astype(np.complex64).astype(np.complex128)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    kw = dict(nlevels=2, n_outer=3, n_cg=4, final_cg=5, accel=1)
    x_ref, T_ref = O.pyramid_solve(y, Sm32, k1, k2, off, N, **kw)
    xg, Tg, info = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, **kw)
    assert info["levels"][1]["outer"] == 0
    assert rel_l2(xg.cpu().numpy(), x_ref) < 1e-4, rel_l2(xg.cpu().numpy(), x_ref)
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01


def test_cg_schedule_follows_solveXT_by_hand():
    """nX of Reconstruction/solveXT.m:417-425 worked by hand for nIt(2) = 8
    and for the reference's default nIt(2) = 1 (solveXT.m:178): levels
    32, 64 (new resolutions) and two repeated full-resolution levels."""
    from disorder_amd import pyramid as PY
    sizes = [[32] * 3, [64] * 3, [64] * 3, [64] * 3]
    want = {
        8: {0: [8, 7, 6, 5, 4, 3, 2, 1, 1, 1], 1: [7, 6, 5, 4, 3, 2, 1, 1], 2: [0, 6, 5, 3, 1, 1],
            3: [0, 6, 5, 3, 1]},
        1: {0: [1, 1, 1], 1: [1, 1, 1], 2: [0, 1, 1, 1], 3: [0, 1]},
        3: {0: [3, 2, 1, 1], 1: [2, 1, 1], 2: [0, 1, 1, 1], 3: [0, 1, 1]},
    }
    for mod in (PY, O):
        for nIt2, per_level in want.items():
            for li, seq in per_level.items():
                got, nX = [], nIt2
                for it in range(1, len(seq) + 1):
                    nX = mod.cg_iterations(li, it, sizes, nX, nIt2)
                    got.append(nX)
                assert got == seq, (mod.__name__, nIt2, li, got, seq)


def test_readout_roi_follows_solveXT():
    """E.nF of solveXT.m:345 (1-based inclusive there, 0-based half-open
    here): 'F' 1:floor((1-redFOV) n3), 'H' floor(1+redFOV n3):n3."""
    from disorder_amd import pyramid as PY
    for mod in (PY, O):
        assert mod.readout_roi(256, "H") == (85, 256)      # 86:256
        assert mod.readout_roi(256, "F") == (0, 170)       # 1:170
        assert mod.readout_roi(64, "H") == (21, 64)        # floor(22.33) = 22 -> 22:64
        assert mod.readout_roi(3, "F") == (0, 2)           # (2/3)*3 rounds up to 2.0000000000000004
        assert mod.readout_roi(3, "H") == (1, 3)
        assert mod.readout_roi(100, "AP") == (0, 100)
        assert mod.readout_roi(90, "H", 0.5) == (45, 90)   # floor(46) = 46 -> 46:90
