"""Repeated acquisitions (NRepeats = size(y, 5) > 1, Reconstruction/solveXT.m:84-141,
Methods/buildHarmonicSampling.m:27-38).

A 16-shot Checkered scan on the full 32 x 24 grid acquired twice (32 shots,
two repeats): every phase-encode line is sampled once per repeat, so each
position appears in two motion states, and the data are [NY1, NY2, N3, C, 2].
One shot of the second repeat moves half-way through its echo train (a
jump of (1.5, -1, 0.8) voxels and (3, -2, 2) degrees), as in the
within-shot tests.
CPU: the oracle driver flags that shot (and at most a few still ones, which
split into equal halves) and splits it at its middle profile; the two
halves' estimates differ by the jump.
GPU: the driver splits the same states and reaches the oracle's image
(NRMSE <= 1e-4) and motion (0.01 voxel, 0.01 deg); the entry point on the
5-D data equals the driver on the gathered samples."""
import math

import numpy as np
import pytest

from disorder_amd import sampling as SA
from disorder_amd import trajectory as TR
from oracle import disorder_oracle as O
from tests.helpers import rel_l2, smooth_volume

JUMP = np.r_[1.5, -1.0, 0.8, 3 * math.pi / 180, -2 * math.pi / 180, 2 * math.pi / 180]
KW = dict(nlevels=2, n_outer=8, n_cg=4, final_cg=6, n_within=2, gibbs=0)
S, R = 16, 2


def _scan(seed=11, N=(32, 24, 16), C=4, bad=16 + 6, noise=0.003):
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(seed)
    K = N[:2]
    g1, g2 = np.meshgrid(np.arange(K[0]) - K[0] // 2, np.arange(K[1]) - K[1] // 2, indexing="ij")
    k = np.stack([g1.ravel(order="F"), g2.ravel(order="F")], 1)
    ko, NP, _, _, _, _ = TR.sampling_disorder(k, (4, 4), "Checkered", "AlternatingZigZag", K=K)
    kTraj = np.concatenate([ko] * R)
    kRange = [[-(K[0] // 2), K[0] - K[0] // 2 - 1], [-(K[1] // 2), K[1] - K[1] // 2 - 1]]
    k1, k2, off, st = TR.sampling_from_trajectory(kTraj, kRange, NP // S, n_repeats=R)
    rep = st["sample_repeat"]
    NS = R * S
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5)
    T = np.zeros((NS, 6))
    T[:, :3] = rng.uniform(-1, 1, (NS, 3))
    T[:, 3:] = rng.uniform(-2, 2, (NS, 3)) * np.pi / 180
    split = np.zeros(NS, bool)
    split[bad] = True
    sop, _ = TR.subdivide_states(st["sweep_of_profile"], split)
    Ti = st["time_index"]
    H = np.where(Ti > 0, sop[np.maximum(Ti, 1) - 1], 0)
    s1, s2, srep, so = SA.harmonic_sampling_repeats(H, NS + 1)
    T2 = np.repeat(T, np.where(split, 2, 1), axis=0)
    T2[bad + 1] = T2[bad] + JUMP
    ysp = O.encode(x_gt, Sm, O.Sampling(s1, s2, so, N[0], N[1]), T2, N)
    pos = {(a, b, r): i for i, (a, b, r) in enumerate(zip(s1, s2, srep))}
    y = ysp[[pos[(a, b, r)] for a, b, r in zip(k1, k2, rep)]]
    sc = np.sqrt(np.mean(np.abs(y) ** 2))
    y = y + noise * sc * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape)) / math.sqrt(2)
    y = y.astype(np.complex64).astype(np.complex128)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    return dict(x_gt=x_gt, Sm=Sm, Sm32=Sm32, y=y, k1=k1, k2=k2, rep=rep, off=off, N=N, st=st, bad=bad,
                kTraj=kTraj, kRange=kRange, echoes=NP // S)


def test_states_and_samples():
    d = _scan()
    st, k1, k2, rep, off = d["st"], d["k1"], d["k2"], d["rep"], d["off"]
    assert st["H"].shape == d["N"][:2] + (R,) and st["sweeps"] == R * S
    # every position once per repeat, repeat r holding the states of its pass
    assert len(k1) == R * d["N"][0] * d["N"][1]
    state = np.repeat(np.arange(R * S), np.diff(off))
    assert np.array_equal(rep, state // S)
    lin = k1 + d["N"][0] * k2
    for r in range(R):
        assert np.array_equal(np.sort(lin[rep == r]), np.arange(d["N"][0] * d["N"][1]))


def test_oracle_splits_the_moving_shot():
    d = _scan()
    info = {}
    _, Tr = O.pyramid_solve(d["y"], d["Sm32"], d["k1"], d["k2"], d["off"], d["N"], traj=d["st"], info=info, **KW)
    sos = info["sweep_of_state"]
    bad = d["bad"]
    split = [s for s in range(R * S) if (sos == s).sum() == 2]
    assert bad in split and len(split) <= 4, split
    for s in split:
        i = np.flatnonzero(sos == s)
        dd = Tr[i[1]] - Tr[i[0]]
        if s == bad:  # the two halves differ by the jump
            assert np.abs(dd[:3] - JUMP[:3]).max() < 0.1, dd
            assert np.abs(dd[3:] - JUMP[3:]).max() < 0.005, dd
        else:  # a still shot flagged with it (shot 6 of the first repeat samples the same lines) splits into equals
            assert np.abs(dd[:3]).max() < 0.05 and np.abs(dd[3:]).max() < 0.002, (s, dd)


@pytest.mark.gpu
def test_driver_split_matches_oracle(dev):
    from disorder_amd import pyramid as PY
    from tests.test_pyramid import _to_dev
    d = _scan()
    info_o = {}
    x_ref, T_ref = O.pyramid_solve(d["y"], d["Sm32"], d["k1"], d["k2"], d["off"], d["N"], traj=d["st"],
                                   info=info_o, **KW)
    xg, Tg, info = PY.solveXT(_to_dev(d["y"], dev), _to_dev(d["Sm"], dev), d["k1"], d["k2"], d["off"], d["N"],
                              traj=d["st"], **KW)
    assert np.array_equal(info["sweep_of_state"], info_o["sweep_of_state"])
    assert (info["sweep_of_state"] == d["bad"]).sum() == 2
    assert rel_l2(xg.cpu().numpy(), x_ref) < 1e-4, rel_l2(xg.cpu().numpy(), x_ref)
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01


@pytest.mark.gpu
def test_entry_point_on_five_dimensional_data(dev):
    """aligned_reconstruction on y [32, 24, 16, 4, 2] (solveXT.m:69, NY(5)
    repeats) gathers row k1 + NY1 k2 + NY1 NY2 r for a sample of repeat r
    (buildHarmonicSampling.m:29-38) and then agrees with the driver on the
    compact samples (motion 1e-4, image NRMSE 1e-5); data of a 4-D array with two passes in one repeat
    slot are refused."""
    from disorder_amd import pyramid as PY
    from disorder_amd import recon
    from tests.test_pyramid import _to_dev
    d = _scan()
    N, k1, k2, rep = d["N"], d["k1"], d["k2"], d["rep"]
    Y = np.zeros(N[:2] + (N[2], d["Sm"].shape[3], R), complex)
    Y[k1, k2, :, :, rep] = d["y"]
    kw = dict(nlevels=2, n_outer=2, n_cg=4, final_cg=4, n_within=1, gibbs=0)
    out = recon.aligned_reconstruction(_to_dev(Y, dev), _to_dev(d["Sm"], dev), d["kTraj"], d["kRange"],
                                       d["echoes"], perc=None, **kw)
    xg, Tg, info = PY.solveXT(_to_dev(d["y"], dev), _to_dev(d["Sm"], dev), k1, k2, d["off"], N, traj=d["st"], **kw)
    # not bitwise: the coil stage of this (non-register) grid scatters samples into LDS rows with float atomics,
    # whose order varies from run to run
    assert np.array_equal(out["sweep_of_state"], info["sweep_of_state"])
    assert np.abs(out["T"] - Tg).max() < 1e-4
    assert rel_l2(out["x"].cpu().numpy(), xg.cpu().numpy()) < 1e-5
    with pytest.raises(ValueError):
        recon.aligned_reconstruction(_to_dev(Y[..., 0], dev), _to_dev(d["Sm"], dev), d["kTraj"], d["kRange"],
                                     d["echoes"], perc=None, **kw)
