"""SENSE folding of the phase encodes (Methods/fold.m, ifold.m,
buildFoldM.m; Reconstruction/encode.m:100-104, decode.m:73-76).

CPU: the oracle's literal fold.m (window slicing, or reshape-sum-circshift
for multiples) equals the literal buildFoldM.m matrix and the alias sum over
r = i + ceil((NX - NY)/2) (mod NY); ifold.m is its adjoint and the
transposed matrix; for NX = 2 NY the DFT of the folded signal is the
NX-point DFT at the even frequencies times exp(2 pi i k o / NY); an object
inside the central NY window folds to its crop, so the folded encode equals
the unfolded encode on the cropped grid; the folded encode/decode pair passes
the adjoint dot test.
GPU: a plan with dis_plan_set_fold against the oracle for encode, decode,
E^H E (weights, motion with dim-1 translation), per-sample energies, CG and
the LM sums, on a non-multiple, an odd-difference and a multiple grid."""
import math

import numpy as np
import pytest

from oracle import disorder_oracle as O
from tests.helpers import crandn, random_motion, rel_l2, smooth_volume

SIZES = [(8, 8), (10, 8), (9, 8), (12, 8), (15, 8), (16, 8), (24, 8), (20, 16), (7, 4), (13, 10), (31, 24),
         (40, 32), (36, 27)]


@pytest.mark.parametrize("NX,NY", SIZES)
def test_fold_matches_matrix_alias_sum_and_adjoint(NX, NY):
    rng = np.random.default_rng(NX * 100 + NY)
    x = crandn(rng, (NX, 3))
    f = O.fold(x, 0, NX, NY)
    A = O.build_fold_matrix(NX, NY)
    assert np.array_equal(A.sum(axis=0), np.ones(NX))          # every sample lands once
    assert np.abs(A @ x - f).max() < 1e-14
    o = math.ceil((NX - NY) / 2)
    ref = np.zeros((NY, 3), complex)
    for r in range(NX):
        ref[(r - o) % NY] += x[r]
    assert np.abs(ref - f).max() < 1e-14
    y = crandn(rng, (NY, 3))
    u = O.ifold(y, 0, NX, NY)
    assert np.abs(A.T @ y - u).max() < 1e-14
    assert abs(np.vdot(y, f) - np.vdot(u, x)) < 1e-12
    # along another axis
    x3 = crandn(rng, (3, NX, 2))
    assert np.abs(O.fold(x3, 1, NX, NY) - np.einsum("ij,ajb->aib", A, x3)).max() < 1e-14


def test_fold_then_dft_is_subsampled_spectrum():
    rng = np.random.default_rng(3)
    NY = 12
    x = crandn(rng, 2 * NY)
    o = math.ceil(NY / 2)
    k = np.arange(NY)
    lhs = np.fft.fft(O.fold(x, 0, 2 * NY, NY))
    rhs = np.fft.fft(x)[2 * k] * np.exp(2j * np.pi * k * o / NY)
    assert np.abs(lhs - rhs).max() < 1e-12


def _fold_problem(rng, n, ny, C=3, S=4, weights=False):
    x = smooth_volume(rng, n) + 0.1 * crandn(rng, n)
    Sm = crandn(rng, n + (C,)) * 0.5
    H = rng.integers(1, S + 1, size=ny)
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    We = rng.uniform(0.5, 1.5, off[S]) if weights else None
    samp = O.Sampling(k1, k2, off[:S + 1], ny[0], ny[1], We)
    return x, Sm, samp


def test_object_inside_window_folds_to_its_crop():
    rng = np.random.default_rng(4)
    n, ny = (20, 18, 6), (16, 12)
    x, Sm, samp = _fold_problem(rng, n, ny)
    o = [math.ceil((n[d] - ny[d]) / 2) for d in range(2)]
    mask = np.zeros(n)
    mask[o[0]:o[0] + ny[0], o[1]:o[1] + ny[1], :] = 1
    x = x * mask
    T0 = np.zeros((samp.S, 6))
    y = O.encode(x, Sm, samp, T0, n)
    crop = (slice(o[0], o[0] + ny[0]), slice(o[1], o[1] + ny[1]))
    yc = O.encode(x[crop], Sm[crop], samp, T0, (ny[0], ny[1], n[2]))
    assert np.abs(y - yc).max() < 1e-12 * np.abs(y).max()


def test_folded_encode_decode_adjoint():
    rng = np.random.default_rng(5)
    n, ny = (20, 15, 6), (16, 12)
    x, Sm, samp = _fold_problem(rng, n, ny, weights=False)
    T = random_motion(rng, samp.S)
    y = crandn(rng, (samp.Ns, n[2], Sm.shape[3]))
    lhs = np.vdot(y, O.encode(x, Sm, samp, T, n))
    rhs = np.vdot(O.decode(y, Sm, samp, T, n), x)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def _to_dev(a, dev):
    import torch
    import disorder_amd as D
    return D.colmajor(torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64))).to(dev))


# (reconstruction grid, acquired phase-encode grid): non-multiple, odd
# difference with a 3^3 acquired length, and multiples (register-path n1);
# folds on register-path n1 run in the register coil kernel: along dim 2
# only (non-multiple, multiple, odd acquired length), along dim 1 by an
# integer factor (alone, and by 3 on the radix-9 length 144 with a
# non-multiple fold along dim 2); the grids with n1 = 40 take the fold stage
GRIDS = [((40, 36, 16), (32, 30)), ((40, 36, 12), (35, 27)), ((64, 48, 16), (32, 24)),
         ((64, 48, 16), (64, 32)), ((64, 48, 12), (64, 24)), ((32, 40, 12), (32, 27)),
         ((64, 48, 12), (32, 48)), ((144, 36, 8), (48, 30)),
         # offset larger than the acquired length (64 -> 18: five-fold aliasing), and a larger grid
         ((32, 64, 8), (32, 18)), ((128, 96, 24), (64, 72)),
         # acquired lengths with prime factors 11 and 13 (runtime-radix FFT stage)
         ((26, 24, 8), (13, 22)), ((32, 30, 11), (22, 26))]


@pytest.mark.gpu
@pytest.mark.parametrize("n,ny", GRIDS)
def test_folded_operators_match_oracle(dev, n, ny):
    import disorder_amd as D
    rng = np.random.default_rng(sum(n) + sum(ny))
    x, Sm, samp = _fold_problem(rng, n, ny, C=3, S=4, weights=True)
    T = random_motion(rng, samp.S)
    Sm = Sm.astype(np.complex64).astype(np.complex128)
    x = x.astype(np.complex64).astype(np.complex128)
    E = D.Encoding(_to_dev(Sm, dev), samp.k1, samp.k2, samp.off, We=samp.We, NY=ny)
    xd = _to_dev(x, dev)
    y_g = E.encode(xd, T).cpu().numpy()
    y_r = O.encode(x, Sm, samp, T, n)
    assert rel_l2(y_g, y_r) < 1e-5, rel_l2(y_g, y_r)
    yn = (y_r + 0.1 * crandn(rng, y_r.shape)).astype(np.complex64).astype(np.complex128)
    yd = _to_dev(yn, dev)
    x_g = E.decode(yd, T).cpu().numpy()
    assert rel_l2(x_g, O.decode(yn, Sm, samp, T, n)) < 1e-5
    ehe = E.encodedecode(xd, T).cpu().numpy()
    assert rel_l2(ehe, O.encodedecode(x, Sm, samp, T, n)) < 1e-5
    e_g = E.energy(xd, yd, T, per_sample=True)
    e_r = O.compute_energy(yn, x, Sm, samp, T, n, per_sample=True)
    assert rel_l2(e_g, e_r) < 1e-5
    # zero motion goes through the coil stage without the transform
    T0 = np.zeros_like(T)
    assert rel_l2(E.encode(xd, T0).cpu().numpy(), O.encode(x, Sm, samp, T0, n)) < 1e-5
    E.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,ny", [((40, 36, 12), (32, 30)), ((64, 48, 12), (64, 30)), ((64, 48, 12), (32, 30))])
def test_folded_cg_and_lm_match_oracle(dev, n, ny):
    import disorder_amd as D
    rng = np.random.default_rng(11)
    x, Sm, samp = _fold_problem(rng, n, ny, C=4, S=4)
    Sm = Sm.astype(np.complex64).astype(np.complex128)
    x = x.astype(np.complex64).astype(np.complex128)
    Tgt = random_motion(rng, samp.S, 3, 1)
    y = O.encode(x, Sm, samp, Tgt, n) + 0.01 * crandn(rng, (samp.Ns, n[2], 4))
    y = y.astype(np.complex64).astype(np.complex128)
    E = D.Encoding(_to_dev(Sm, dev), samp.k1, samp.k2, samp.off, NY=ny)
    yd = _to_dev(y, dev)
    xg, _, _ = D.CGsolver(yd, E, Tgt, nIt=6, tolType="RelativeResidual2Error", tol=0.0)
    xr = O.cg_solver(y, Sm, samp, Tgt, n, nIt=6, tolType="RelativeResidual2Error", tol=0.0)[0]
    assert rel_l2(xg.cpu().numpy(), xr) < 1e-4, rel_l2(xg.cpu().numpy(), xr)
    T0 = Tgt + random_motion(rng, samp.S, 1.0, 0.3)
    En, dG, dH = D.lm_partials(yd, E, _to_dev(x, dev), T0)
    Er, dGr, dHr = O.lm_partials(y, x, Sm, samp, T0, n)
    assert rel_l2(En, Er) < 1e-5 and rel_l2(dG, dGr) < 1e-4 and rel_l2(dH, dHr) < 1e-4
    E.close()



def _folded_scan(seed=31, N=(40, 30, 16), NY=(32, 24), S=16, C=4):
    from disorder_amd import sampling as SA
    from tests.helpers import checkered_sampling
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(seed)
    x = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5).astype(np.complex64).astype(np.complex128)
    H = checkered_sampling(rng, NY[0], NY[1], S, 4, 4)
    k1, k2, off = SA.harmonic_sampling(H, S)
    T = np.zeros((S, 6))
    T[:, :3] = rng.uniform(-1, 1, (S, 3))
    T[:, 3:] = rng.uniform(-2, 2, (S, 3)) * np.pi / 180
    y = O.encode(x, Sm, O.Sampling(k1, k2, off, NY[0], NY[1]), T, N).astype(np.complex64).astype(np.complex128)
    return x, Sm, y, k1, k2, off, N, NY, T


def test_oracle_driver_fold_identity_and_motion():
    """NY = N leaves the oracle driver unchanged; on folded data its motion
    estimates fit the data better than a motion-free reconstruction."""
    from disorder_amd import sampling as SA
    from tests.helpers import checkered_sampling
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(3)
    N, S = (16, 12, 8), 4
    x = smooth_volume(rng, N)
    Sm = _coil_maps_np(N, 3, 5)
    k1, k2, off = SA.harmonic_sampling(checkered_sampling(rng, N[0], N[1], S, 2, 2), S)
    y = O.encode(x, Sm, O.Sampling(k1, k2, off, N[0], N[1]), random_motion(rng, S, 2, 0.5), N)
    a = O.pyramid_solve(y, Sm, k1, k2, off, N, nlevels=2, n_outer=2, n_cg=3)
    b = O.pyramid_solve(y, Sm, k1, k2, off, N, nlevels=2, n_outer=2, n_cg=3, NY=N[:2])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    x_gt, Sm, y, k1, k2, off, N, NY, _ = _folded_scan()
    samp = O.Sampling(k1, k2, off, NY[0], NY[1])
    x_r, T_r = O.pyramid_solve(y, Sm, k1, k2, off, N, nlevels=2, n_outer=3, n_cg=4, NY=NY, gibbs=0)
    T0 = np.zeros_like(T_r)
    x_0 = O.cg_solver(y, Sm, samp, T0, N, nIt=12, tolType="RelativeResidual2Error", tol=0.0)[0]
    e_motion = O.compute_energy(y, x_r, Sm, samp, T_r, N)
    e_still = O.compute_energy(y, x_0, Sm, samp, T0, N)
    assert e_motion < 0.5 * e_still, (e_motion, e_still)


@pytest.mark.gpu
@pytest.mark.parametrize("N,NY", [((40, 30, 16), (32, 24)), ((64, 48, 16), (32, 36))])
def test_folded_driver_matches_oracle(dev, N, NY):
    """The pyramid driver on folded data (two levels; 40x30x16 on 32x24
    through the fold kernels, 64x48x16 on 32x36 in the register coil kernel
    at both levels): GPU against the FP64 oracle driver, image NRMSE <= 1e-4,
    motion within 0.01 voxel and 0.01 degree."""
    from disorder_amd import pyramid as PY
    x_gt, Sm, y, k1, k2, off, N, NY, _ = _folded_scan(N=N, NY=NY)
    x_ref, T_ref = O.pyramid_solve(y, Sm, k1, k2, off, N, nlevels=2, n_outer=2, n_cg=4, NY=NY)
    xg, Tg, info = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, nlevels=2, n_outer=2, n_cg=4,
                              NY=NY)
    assert rel_l2(xg.cpu().numpy(), x_ref) < 1e-4, rel_l2(xg.cpu().numpy(), x_ref)
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01


@pytest.mark.gpu
def test_folded_end_to_end_matches_driver(dev):
    """aligned_reconstruction on data acquired on a 32x24 grid with 40x36
    maps: the same image and motion as the driver called with NY on the
    compact samples (the entry point only gathers and forwards); with cs the
    shearlet-regularised image stays close to the unregularised one."""
    from disorder_amd import pyramid as PY
    from disorder_amd import recon
    from disorder_amd import trajectory as TR
    from tests.test_pyramid import _coil_maps_np
    N, NY, S, C = (40, 36, 35), (32, 24), 16, 3  # every shearlet plane >= 33
    rng = np.random.default_rng(9)
    g1, g2 = np.meshgrid(np.arange(NY[0]) - NY[0] // 2, np.arange(NY[1]) - NY[1] // 2, indexing="ij")
    k = np.stack([g1.ravel(order="F"), g2.ravel(order="F")], 1)
    ko, NP, _, _, _, _ = TR.sampling_disorder(k, (4, 4), "Checkered", "AlternatingZigZag", K=NY)
    kRange = [[-(NY[0] // 2), NY[0] - NY[0] // 2 - 1], [-(NY[1] // 2), NY[1] - NY[1] // 2 - 1]]
    k1, k2, off, st = TR.sampling_from_trajectory(ko, kRange, NP // S)
    x = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5)
    y = O.encode(x, Sm, O.Sampling(k1, k2, off, NY[0], NY[1]), random_motion(rng, S, 2, 1), N)
    Y = np.zeros(NY + (N[2], C), complex)
    Y[k1, k2] = y
    kw = dict(nlevels=2, n_outer=2, n_cg=4, final_cg=4)
    out = recon.aligned_reconstruction(_to_dev(Y, dev), _to_dev(Sm, dev), ko, kRange, NP // S, perc=None, **kw)
    xg, Tg, _ = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, traj=st, NY=NY, **kw)
    assert np.array_equal(out["T"], Tg)
    assert rel_l2(out["x"].cpu().numpy(), xg.cpu().numpy()) < 1e-6
    # the shearlet-regularised image on folded data (folded fidelity trace)
    o2 = recon.aligned_reconstruction(_to_dev(Y, dev), _to_dev(Sm, dev), ko, kRange, NP // S, perc=None, cs=True, **kw)
    xc = o2["x_cs"].cpu().numpy()
    assert xc.shape == N and np.isfinite(xc).all() and len(o2["lambdas"]) == 2
    assert rel_l2(xc, out["x"].cpu().numpy()) < 0.5


def test_oracle_folded_fidelity_sample():
    """traceEstimation's fidelity sample with folding: unit weights give
    sum_c ||fold(S_c x)||^2 (Parseval on the acquired grid); NY = NX is the
    unfolded formula."""
    from oracle import shearlet_oracle as SO
    rng = np.random.default_rng(6)
    n, NY = (20, 15, 6), (16, 12)
    Sm = crandn(rng, n + (3,))
    z = rng.normal(size=n)
    got = SO.trace_samples_fidelity([z], Sm, np.ones(NY))[0]
    ref = sum(float(np.sum(np.abs(O.fold(O.fold(Sm[..., c] * z, 0, n[0], NY[0]), 1, n[1], NY[1])) ** 2))
              for c in range(3))
    assert abs(got / ref - 1) < 1e-12
    W = rng.uniform(0.2, 1.0, n[:2])
    un = sum(float(np.sum(W[:, :, None] * np.abs(np.fft.fft2(Sm[..., c] * z, axes=(0, 1))) ** 2)) / (n[0] * n[1])
             for c in range(3))
    assert abs(SO.trace_samples_fidelity([z], Sm, W)[0] / un - 1) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("n,ny", [((40, 36, 12), (32, 30)), ((64, 48, 16), (32, 24))])
def test_folded_fidelity_trace_matches_oracle(dev, n, ny):
    import torch
    from disorder_amd import shearlet as SH
    from oracle import shearlet_oracle as SO
    import disorder_amd as D
    rng = np.random.default_rng(sum(n))
    Sm = crandn(rng, n + (3,)).astype(np.complex64).astype(np.complex128)
    z = rng.normal(size=n).astype(np.float32).astype(np.float64)
    W = rng.uniform(0.2, 1.0, ny).astype(np.float32)
    Wd = torch.from_numpy(W.reshape(-1, order="F")).to(dev)
    got = SH.fidelity_folded(_to_dev(Sm, dev), Wd, _to_dev(z, dev), ny, D.default_context(0))
    ref = SO.trace_samples_fidelity([z], Sm, W.astype(np.float64))[0]
    assert abs(got / ref - 1) < 2e-5, (got, ref)
