"""End to end (disorder_amd.recon.aligned_reconstruction): a DISORDER scan
from its trajectory (Checkered order on an elliptical shutter, 16 shots),
6 coils compressed, rigid motion per shot (zero mean, as
Simulation/synthesizeT.m draws it), data on the full phase-encode grid.
The corrected image is much closer to the truth than the reconstruction
that ignores motion, the motion is recovered, the CS image exists, and the
NIfTI/MAT outputs read back with the right shapes."""
import math
import os

import numpy as np
import pytest

from disorder_amd import trajectory as TR
from oracle import disorder_oracle as O
from tests.golden.make_trajectory_fixture import shutter_points
from tests.helpers import rel_l2, smooth_volume


@pytest.mark.gpu
def test_aligned_reconstruction_end_to_end(dev, tmp_path):
    import torch
    import disorder_amd as D
    from disorder_amd import nifti
    from disorder_amd import recon
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(40)
    N, C, S = (40, 36, 35), 6, 16  # every plane >= 33 for the CS shearlets
    k = shutter_points(N[0], N[1])
    k = k[: len(k) - len(k) % S]
    ko, NP, _, _, _, _ = TR.sampling_disorder(k, (4, 4), "Checkered", "AlternatingZigZag", K=N[:2])
    kRange = [[-(N[0] // 2), N[0] - N[0] // 2 - 1], [-(N[1] // 2), N[1] - N[1] // 2 - 1]]
    k1, k2, off, st = TR.sampling_from_trajectory(ko, kRange, NP // S)
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5)
    T_gt = np.zeros((S, 6))
    T_gt[:, :3] = rng.uniform(-1.5, 1.5, (S, 3))
    T_gt[:, 3:] = rng.uniform(-3, 3, (S, 3)) * math.pi / 180
    T_gt -= T_gt.mean(axis=0)
    yc = O.encode(x_gt, Sm, O.Sampling(k1, k2, off, N[0], N[1]), T_gt, N)
    yfull = np.zeros(N[:2] + (N[2], C), np.complex128)
    yfull[k1, k2] = yc
    to = lambda a: D.colmajor(torch.from_numpy(np.asfortranarray(a.astype(np.complex64))).to(dev))
    out = recon.aligned_reconstruction(to(yfull), to(Sm), ko, kRange, NP // S, perc=(0.98, 0.98, 0.98), nlevels=2,
                                       n_outer=8, n_cg=6, final_cg=12, n_within=1, cs=True, uncorrected_cg=12,
                                       write=(str(tmp_path), "scan"))
    x = out["x"].cpu().numpy()
    xu = out["x_uncorrected"].cpu().numpy()
    e_c, e_u = rel_l2(x, x_gt), rel_l2(xu, x_gt)
    print(f"\n[end-to-end] eivaS {out['eivaS']}; NRMSE corrected {e_c:.4f} uncorrected {e_u:.4f}")
    assert e_c < 0.5 * e_u, (e_c, e_u)
    T = out["T"]
    sos = out["sweep_of_state"]
    Tm = np.stack([T[sos == s].mean(axis=0) for s in range(S)])
    assert np.abs(Tm[:, :3] - T_gt[:, :3]).max() < 0.3
    assert np.abs(Tm[:, 3:] - T_gt[:, 3:]).max() * 180 / math.pi < 0.75  # of +-3 degrees, 4 coils, 2 levels
    assert np.isfinite(out["x_cs"].cpu().numpy()).all() and len(out["lambdas"]) == 2
    names = sorted(os.path.basename(p) for p in out["files"])
    assert names == sorted(["scan_Di.nii", "scan_DiPh.nii", "scan_Aq.nii", "scan_AqPh.nii", "scan_Re.nii",
                            "scan_RePh.nii", "scan_Tr.mat"])
    back, _, _ = nifti.read_pair(os.path.join(str(tmp_path), "An-Ve", "scan"), "Di")
    assert back.shape == N and rel_l2(back, x) < 1e-6


def test_entry_point_checks_data_before_touching_the_device():
    """Shape, coil count and memory layout of y are checked before any
    device work: a row-major 4-D y (read through a column-major view below)
    and a coil count that differs from the maps' are refused."""
    import torch
    from disorder_amd import colmajor, recon
    S = colmajor(torch.zeros((8, 6, 4, 2), dtype=torch.complex64))
    kTraj, kRange = np.zeros((2, 2), np.int64), [[-4, 3], [-3, 2]]
    with pytest.raises(ValueError, match="column-major"):
        recon.aligned_reconstruction(torch.zeros((8, 6, 4, 2), dtype=torch.complex64), S, kTraj, kRange, 1)
    with pytest.raises(ValueError, match="same coils"):
        recon.aligned_reconstruction(colmajor(torch.zeros((8, 6, 4, 3), dtype=torch.complex64)), S, kTraj, kRange, 1)
    with pytest.raises(ValueError, match="NY1"):
        recon.aligned_reconstruction(colmajor(torch.zeros((8, 6, 4), dtype=torch.complex64)), S, kTraj, kRange, 1)


def _small_scan(N=(16, 12, 6), S=4):
    K = N[:2]
    g1, g2 = np.meshgrid(np.arange(K[0]) - K[0] // 2, np.arange(K[1]) - K[1] // 2, indexing="ij")
    k = np.stack([g1.ravel(order="F"), g2.ravel(order="F")], 1)
    kRange = [[-(K[0] // 2), K[0] - K[0] // 2 - 1], [-(K[1] // 2), K[1] - K[1] // 2 - 1]]
    return k, kRange


def test_reference_early_exits_return_fail_before_device_work():
    """solveXT.m:33 (too many dynamics), :77 (separable y-z trajectory) and
    :83 (interrupted scan) set rec.Fail and return: the entry point returns
    {"fail": reason} from host checks alone (CPU tensors here: nothing
    touches a device)."""
    import torch
    from disorder_amd import recon
    N = (16, 12, 6)
    k, kRange = _small_scan(N)
    S = torch.zeros(N + (2,), dtype=torch.complex64).permute(3, 2, 1, 0).contiguous().permute(3, 2, 1, 0)
    y4 = torch.zeros(N[:2] + (N[2], 2), dtype=torch.complex64).permute(3, 2, 1, 0).contiguous().permute(3, 2, 1, 0)
    y5 = torch.zeros(N[:2] + (N[2], 2, 4), dtype=torch.complex64)
    out = recon.aligned_reconstruction(y5, S, k, kRange, len(k) // 4)
    assert "too big" in out["fail"] and "4 dynamics" in out["fail"]
    out = recon.aligned_reconstruction(y4, S, (k[:, 0], k[:-3, 1]), kRange, 1)
    assert "SEPARABLE" in out["fail"]
    out = recon.aligned_reconstruction(y4, S, k[:1], kRange, 1)
    assert "SEPARABLE" in out["fail"]
    out = recon.aligned_reconstruction(y4, S, k[:-5], kRange, 16)
    assert "INTERRUPTED" in out["fail"]


def test_trajectory_pair_equals_stacked_array():
    k, kRange = _small_scan()
    a = TR.trajectory_states(k, kRange, 48)
    b = TR.trajectory_states((k[:, 0].tolist(), k[:, 1].tolist()), kRange, 48)
    assert np.array_equal(a["H"], b["H"]) and np.array_equal(a["sweep_of_profile"], b["sweep_of_profile"])
    with pytest.raises(TR.SeparableTrajectory):
        TR.trajectory_states((k[:, 0], k[:-1, 1]), kRange, 48)
    with pytest.raises(TR.InterruptedScan):
        TR.trajectory_states(k[:-1], kRange, 48)


@pytest.mark.gpu
def test_not_disorder_returns_uncorrected_reconstruction(dev):
    """Sequential shots (each shot a block of phase-encode rows) leave the
    outer shots without samples in the coarse level's central k-space:
    solveXT.m:355 returns without corrections.  The driver raises
    NotDisorder; the entry point returns nowithin with the reconstruction
    without motion correction, which equals a plain CG at T = 0."""
    import torch
    import disorder_amd as D
    from disorder_amd import pyramid as PY, recon
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(3)
    N, C, S = (32, 24, 8), 3, 4
    K = N[:2]
    # sequential order: shot s takes phase-encode rows k1 in [8 s - 16, 8 s - 8) (all k2)
    g1, g2 = np.meshgrid(np.arange(K[0]) - K[0] // 2, np.arange(K[1]) - K[1] // 2, indexing="ij")
    k = np.stack([g1.ravel(order="C"), g2.ravel(order="C")], 1)
    kRange = [[-(K[0] // 2), K[0] - K[0] // 2 - 1], [-(K[1] // 2), K[1] - K[1] // 2 - 1]]
    k1, k2, off, st = TR.sampling_from_trajectory(k, kRange, len(k) // S)
    x = smooth_volume(rng, N, nblob=6)
    Sm = _coil_maps_np(N, C, 4)
    yc = O.encode(x, Sm, O.Sampling(k1, k2, off, N[0], N[1]), np.zeros((S, 6)), N)
    to = lambda a: D.colmajor(torch.from_numpy(np.asfortranarray(a.astype(np.complex64))).to(dev))
    with pytest.raises(PY.NotDisorder):
        PY.solveXT(to(yc), to(Sm), k1, k2, off, N, nlevels=3, n_outer=2, n_cg=2, gibbs=0)
    yfull = np.zeros(N[:2] + (N[2], C), np.complex128)
    yfull[k1, k2] = yc
    out = recon.aligned_reconstruction(to(yfull), to(Sm), k, kRange, len(k) // S, perc=None, nlevels=3, n_outer=2,
                                       n_cg=2, final_cg=10, gibbs=0)
    assert out["nowithin"] and "not DISORDER" in out["message"] and np.all(out["T"] == 0)
    E = D.Encoding(to(Sm), k1, k2, off)
    xr, _, _ = D.CGsolver(to(yc), E, np.zeros((S, 6)), nIt=10, tolType="Energy", tol=1e-3,
                          Se=D.coil_precond(to(Sm)))
    assert rel_l2(out["x"].cpu().numpy(), xr.cpu().numpy()) < 1e-5
