"""Within-shot subdivision (Reconstruction/solveXT.m:270-287, :470-478).

A 16-shot DISORDER scan (Checkered order, full 32 x 24 grid) in which one
shot moves half-way through its echo train: its first half is encoded with
the shot's motion, its second half with that motion plus a jump of
(1.5, -1, 0.8) voxels and (3, -2, 2) degrees.
CPU: in the oracle driver the first full-resolution level flags that shot
alone, the next level splits it at its middle profile, and the difference
of the two halves' estimates recovers the jump (0.1 voxel, 0.005 rad).
GPU: the driver makes the same split and reaches the oracle's image
(NRMSE <= 1e-4) and motion (0.01 voxel, 0.01 deg)."""
import math

import numpy as np
import pytest

from disorder_amd import sampling as SA
from disorder_amd import trajectory as TR
from oracle import disorder_oracle as O
from tests.helpers import rel_l2, smooth_volume

JUMP = np.r_[1.5, -1.0, 0.8, 3 * math.pi / 180, -2 * math.pi / 180, 2 * math.pi / 180]
# gibbs = 0: the Tukey window would zero the data at the corners of this full
# rectangular grid and flag the shots that sample them as well.  The fixed
# CG count keeps the image converged enough at the full-resolution level
# that only the moving shot is an outlier; under the nX schedule of
# solveXT.m:417-425 (4, 3, 2, 1, 1, ... CG iterations) the weights flag
# further shots, which the GPU parity test below covers as well.
KW = dict(nlevels=2, n_outer=8, n_cg=4, final_cg=6, n_within=2, gibbs=0, cg_schedule="fixed")


def _scan(seed=5, N=(32, 24, 16), C=4, bad=6, noise=0.003):
    from tests.test_pyramid import _coil_maps_np
    rng = np.random.default_rng(seed)
    K = N[:2]
    g1, g2 = np.meshgrid(np.arange(K[0]) - K[0] // 2, np.arange(K[1]) - K[1] // 2, indexing="ij")
    k = np.stack([g1.ravel(order="F"), g2.ravel(order="F")], 1)
    ko, NP, _, _, _, _ = TR.sampling_disorder(k, (4, 4), "Checkered", "AlternatingZigZag", K=K)
    kRange = [[-(K[0] // 2), K[0] - K[0] // 2 - 1], [-(K[1] // 2), K[1] - K[1] // 2 - 1]]
    S = 16
    k1, k2, off, st = TR.sampling_from_trajectory(ko, kRange, NP // S)
    x_gt = smooth_volume(rng, N, nblob=8)
    Sm = _coil_maps_np(N, C, 5)
    T = np.zeros((S, 6))
    T[:, :3] = rng.uniform(-1, 1, (S, 3))
    T[:, 3:] = rng.uniform(-2, 2, (S, 3)) * np.pi / 180
    # encode the moving shot as two states split at its middle profile
    split = np.zeros(S, bool)
    split[bad] = True
    sop, _ = TR.subdivide_states(st["sweep_of_profile"], split)
    Ti = st["time_index"]
    H = np.where(Ti > 0, sop[np.maximum(Ti, 1) - 1], 0)
    s1, s2, so = SA.harmonic_sampling(H, S + 1)
    T2 = np.repeat(T, np.where(split, 2, 1), axis=0)
    T2[bad + 1] = T2[bad] + JUMP
    ysp = O.encode(x_gt, Sm, O.Sampling(s1, s2, so, N[0], N[1]), T2, N)
    pos = {(a, b): r for r, (a, b) in enumerate(zip(s1, s2))}
    y = ysp[[pos[(a, b)] for a, b in zip(k1, k2)]]
    sc = np.sqrt(np.mean(np.abs(y) ** 2))
    y = y + noise * sc * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape)) / math.sqrt(2)
    y = y.astype(np.complex64).astype(np.complex128)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    return x_gt, Sm, Sm32, y, k1, k2, off, N, st, bad


def test_oracle_splits_the_moving_shot():
    x_gt, _, Sm32, y, k1, k2, off, N, st, bad = _scan()
    info = {}
    _, Tr = O.pyramid_solve(y, Sm32, k1, k2, off, N, traj=st, info=info, **KW)
    sos = info["sweep_of_state"]
    assert sos.tolist() == sorted(list(range(16)) + [bad])
    i = np.flatnonzero(sos == bad)
    d = Tr[i[1]] - Tr[i[0]]
    assert np.abs(d[:3] - JUMP[:3]).max() < 0.1, d
    assert np.abs(d[3:] - JUMP[3:]).max() < 0.005, d


@pytest.mark.gpu
@pytest.mark.parametrize("schedule,N", [("fixed", (32, 24, 16)), ("nX", (64, 24, 16))])
def test_driver_split_matches_oracle(dev, schedule, N):
    """Same split, image and motion as the oracle driver.  With the nX
    schedule the within-shot levels start with no CG (nX = 0,
    solveXT.m:422) and run 1-CG outer iterations, where an LM accept test
    between energies equal to FP32 rounding can go either way; the scan is
    then 64 x 24 x 16, whose levels (32 and 64 along dim 1) both run the
    deterministic register coil path, and the oracle's robust-weight
    decisions there sit 0.6 decades or more from the threshold."""
    from disorder_amd import pyramid as PY
    from tests.test_pyramid import _to_dev
    x_gt, Sm, Sm32, y, k1, k2, off, N, st, bad = _scan(N=N)
    kw = dict(KW, cg_schedule=schedule)
    info_o = {}
    x_ref, T_ref = O.pyramid_solve(y, Sm32, k1, k2, off, N, traj=st, info=info_o, **kw)
    xg, Tg, info = PY.solveXT(_to_dev(y, dev), _to_dev(Sm, dev), k1, k2, off, N, traj=st, **kw)
    assert np.array_equal(info["sweep_of_state"], info_o["sweep_of_state"])
    assert [lv["states"] for lv in info["levels"]] == [16, 16, 16, 17]
    assert rel_l2(xg.cpu().numpy(), x_ref) < 1e-4, rel_l2(xg.cpu().numpy(), x_ref)
    assert np.abs(Tg[:, :3] - T_ref[:, :3]).max() < 0.01
    assert np.abs(Tg[:, 3:] - T_ref[:, 3:]).max() * 180 / np.pi < 0.01
