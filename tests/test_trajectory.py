"""Scanner trajectories (Acquisition/samplingDISORDER.m,
Acquisition/electrostaticRepulsionDISORDER.m, Reconstruction/solveXT.m:75-141,
:270-287, :349-350).  All integer and bit-exact, CPU only.

Pins: hand-worked electrostatic orders for 2x2 and 3x3 tiles; closed
properties of the orders on full grids and on an elliptical shutter (every
shot takes one point of every inside tile, the output is a permutation of
the input, the zig-zag keys are monotone along each shot, alternating
direction on even shots); the state map of a shot-ordered trajectory equals
the shot map that the order reports; a triangle-wave steady-state
trajectory of P periods splits into 2P sweeps; a hand-worked subdivision.
The product (vectorised) equals the oracle's literal loops on every segment
and sample order with the same seeded permutations, and the deterministic
'Checkered' order equals the committed fixture."""
import math
import os

import numpy as np
import pytest

from disorder_amd import sampling as SA
from disorder_amd import trajectory as TR
from oracle import trajectory_oracle as TO
from tests.golden.make_trajectory_fixture import CASES, shutter_points

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "trajectory_checkered.npz")


@pytest.mark.parametrize("impl", [TO.electrostatic_order, TR.electrostatic_order])
def test_electrostatic_hand_orders(impl):
    # 2x2: start at (1,1); (2,2) is farthest on the torus; then ties keep the lower index
    assert list(impl([2, 2])) == [1, 4, 2, 3]
    # 3x3: centre (2,2) = 5, a corner, the opposite corner, then the tie/greedy sequence
    assert list(impl([3, 3])) == [5, 1, 9, 2, 4, 3, 7, 6, 8]


@pytest.mark.parametrize("U", [(4, 4), (5, 3), (2, 6), (8, 8)])
def test_electrostatic_product_matches_oracle(U):
    a = TO.electrostatic_order(U)
    b = TR.electrostatic_order(U)
    assert np.array_equal(a, b)
    assert sorted(b) == list(range(1, U[0] * U[1] + 1))


def test_checkered_matches_fixture():
    z = np.load(GOLDEN)
    for i, (K, U, so) in enumerate(CASES):
        ko, NP, Ko, _, shot, echo = TR.sampling_disorder(z[f"k_in_{i}"], U, "Checkered", so, K=K)
        assert np.array_equal(ko, z[f"k_out_{i}"]), i
        assert np.array_equal(shot, z[f"shot_{i}"]) and np.array_equal(echo, z[f"echo_{i}"]), i


@pytest.mark.parametrize("seg", TR.SEGMENT_ORDERS)
@pytest.mark.parametrize("so", TR.SAMPLE_ORDERS)
def test_product_matches_oracle_all_orders(seg, so):
    for K, U in (((24, 20), (4, 4)), ((18, 16), (3, 2))):
        k = shutter_points(*K)
        S = U[0] * U[1]
        k = k[: len(k) - len(k) % S]
        a = TO.sampling_disorder(k, U, seg, so, K=K, rng=np.random.default_rng(11))
        b = TR.sampling_disorder(k, U, seg, so, K=K, rng=np.random.default_rng(11))
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])


def _zigzag_keys(kout, U, K, S):
    kg = [np.arange(1, v + 1) - (v + 2) // 2 for v in K]
    Ul = [TR._tiles(kg[d], U[d]) for d in range(2)]
    m = TR._tile_of(kout[:, 0], Ul[0])
    n = TR._tile_of(kout[:, 1], Ul[1])
    even = m % 2 == 0
    return 1e9 * m + 1e6 * np.where(even, Ul[1].shape[1] + 1 - n, n) + 1e3 * kout[:, 0] + \
        np.where(even, -kout[:, 1], kout[:, 1])


@pytest.mark.parametrize("so", TR.SAMPLE_ORDERS)
def test_checkered_full_grid_properties(so):
    K, U = (16, 12), (4, 4)
    g1, g2 = np.meshgrid(np.arange(K[0]) - K[0] // 2, np.arange(K[1]) - K[1] // 2, indexing="ij")
    k = np.stack([g1.ravel(order="F"), g2.ravel(order="F")], 1)
    S = 16
    ko, NP, _, _, shot, echo = TR.sampling_disorder(k, U, "Checkered", so, K=K)
    # a permutation of the input
    assert sorted(map(tuple, ko)) == sorted(map(tuple, k))
    NE = NP // S
    for s in range(S):
        blk = ko[s * NE:(s + 1) * NE]
        # one point of every 4x4 tile (tiles start at multiples of 4 offset by ceil(U/2) = 2)
        t = ((blk[:, 0] - 2) // 4, (blk[:, 1] - 2) // 4)
        assert len(set(zip(*t))) == NE
        # zig-zag keys monotone, reversed on even (1-based) shots
        key = _zigzag_keys(blk, U, K, S)
        d = np.diff(key)
        assert (d > 0).all() if (so == "ZigZag" or s % 2 == 0) else (d < 0).all()
        # the maps name the shot and echo of every point
        pos = blk - np.array([-(K[0] // 2), -(K[1] // 2)])
        assert (shot[pos[:, 0], pos[:, 1]] == s + 1).all()
        assert (echo[pos[:, 0], pos[:, 1]] == np.arange(1, NE + 1)).all()
    # electrostatic placement: the j-th point of a tile (column-major) goes to
    # slot order[j] of its block of S, and slot s of every block is shot s
    order = TR.electrostatic_order(U)
    for j, sh in enumerate(order):
        a, b = j % 4, j // 4
        assert shot[-6 + a + K[0] // 2, -6 + b + K[1] // 2] == sh


def test_shutter_tiles_one_point_per_shot():
    K, U = (32, 32), (4, 4)
    z = np.load(GOLDEN)
    k, ko, shot = z["k_in_0"], z["k_out_0"], z["shot_0"]
    assert sorted(map(tuple, ko)) == sorted(map(tuple, k))
    # every 4x4 tile whose corners lie in the shutter holds shots 1..16 once each
    full = 0
    for a in range(-14, 14, 4):
        for b in range(-14, 14, 4):
            corners = [(a, b), (a + 3, b), (a, b + 3), (a + 3, b + 3)]
            if all((c[0] / 16) ** 2 + (c[1] / 16) ** 2 <= 1 for c in corners):
                blk = shot[a + 16:a + 20, b + 16:b + 20]
                assert sorted(blk.ravel()) == list(range(1, 17))
                full += 1
    assert full > 20


def test_trajectory_states_match_shot_map():
    K, U = (24, 20), (4, 4)
    k = shutter_points(*K)
    S = 16
    k = k[: len(k) - len(k) % S]
    ko, NP, _, _, shot, _ = TR.sampling_disorder(k, U, "RandomCheckered", "AlternatingZigZag", K=K,
                                                 rng=np.random.default_rng(2))
    kRange = [[-(K[0] // 2), K[0] - K[0] // 2 - 1], [-(K[1] // 2), K[1] - K[1] // 2 - 1]]
    st = TR.trajectory_states(ko, kRange, NP // S)
    assert st["shots"] == S and st["sweeps"] == S and not st["steady_state"]
    assert np.array_equal(st["sweep_of_profile"], np.repeat(np.arange(1, S + 1), NP // S))
    assert np.array_equal(st["H"], np.fft.ifftshift(shot))
    ref = TO.trajectory_states(ko, kRange, NP // S)
    for key in ("H", "time_index", "sweep_of_profile"):
        assert np.array_equal(st[key], ref[key]), key
    # the state map feeds the harmonic sampling: shot s owns its NP/S lines
    k1, k2, off, _ = TR.sampling_from_trajectory(ko, kRange, NP // S)
    assert np.array_equal(np.diff(off), np.full(S, NP // S))
    assert (st["H"][k1, k2] == np.repeat(np.arange(1, S + 1), NP // S)).all()
    with pytest.raises(ValueError):
        TR.trajectory_states(ko[:-1], kRange, NP // S)
    with pytest.raises(ValueError):
        TR.trajectory_states(np.concatenate([ko, ko[:NP // S]]), kRange, NP // S)


@pytest.mark.parametrize("P", [2, 3, 5])
def test_steady_state_sweeps(P):
    """One echo train (steady state), 2P half-sweeps of M profiles: k1 a
    triangle wave of P periods, k2 = c k1 + p on half-sweep p (c = 2P keeps
    every position distinct).  The dominant orthonormal DCT-II frequency of
    both columns is 2P (index 2P + 1), so the scan splits into 2P sweeps of
    M profiles."""
    M, c = 12, 2 * P
    up = np.arange(M) - M // 2
    tri = np.concatenate([up, up[::-1]] * P)
    half = np.repeat(np.arange(2 * P), M)
    kTraj = np.stack([tri, c * tri + half], 1)
    kRange = [[-M // 2, M // 2 - 1], [int(kTraj[:, 1].min()), int(kTraj[:, 1].max())]]
    A = TO.dct2_ortho(kTraj.astype(float))
    assert (np.argmax(np.abs(A[:len(tri) // 2]), axis=0) == 2 * P).all()
    a = TR.trajectory_states(kTraj, kRange, 1)
    b = TO.trajectory_states(kTraj, kRange, 1)
    assert a["steady_state"] and a["sweeps"] == b["sweeps"] == 2 * P
    assert np.array_equal(a["sweep_of_profile"], np.repeat(np.arange(1, 2 * P + 1), M))
    assert np.array_equal(a["H"], b["H"]) and np.array_equal(a["time_index"], b["time_index"])
    from scipy.fft import dct
    assert np.allclose(A, dct(kTraj.astype(float), type=2, norm="ortho", axis=0), atol=1e-10)


def test_subdivide_states_hand_example():
    st = np.array([1, 1, 1, 1, 2, 2, 2, 3, 3])
    for impl in (TR.subdivide_states, TO.subdivide_states):
        r = impl(st, np.array([True, False, True]) if impl is TR.subdivide_states else [2, 1, 2])
        r = r[0] if isinstance(r, tuple) else r
        assert r.tolist() == [1, 1, 2, 2, 3, 3, 3, 4, 5]
    new, new_of_old = TR.subdivide_states(st, [True, False, True])
    assert new_of_old.tolist() == [1, 3, 4]
    # a one-profile state moves whole to the next label, leaving its own empty
    assert TR.subdivide_states([1, 2, 2], [True, False])[0].tolist() == [2, 3, 3]


def test_repeats_multishot_slots():
    """Two passes over the shutter in 2 x 16 shots, two repeats (solveXT.m:84-141):
    no repetition of the list (32 shots >= 2 repeats), not steady state
    (32 / 2 != 1), 32 sweeps; every position's first acquisition (shots
    1..16) takes repeat slot 1 and its second (shots 17..32) slot 2, so the
    state map is the shot map in slot 1 and the shot map + 16 in slot 2.
    The harmonic sampling returns each state's lines with their repeat, and
    a third acquisition of a position is refused (the reference drops it)."""
    K, U = (24, 20), (4, 4)
    k = shutter_points(*K)
    S = 16
    k = k[: len(k) - len(k) % S]
    ko, NP, _, _, shot, _ = TR.sampling_disorder(k, U, "RandomCheckered", "AlternatingZigZag", K=K,
                                                 rng=np.random.default_rng(5))
    kRange = [[-(K[0] // 2), K[0] - K[0] // 2 - 1], [-(K[1] // 2), K[1] - K[1] // 2 - 1]]
    k2p = np.concatenate([ko, ko])
    st = TR.trajectory_states(k2p, kRange, NP // S, n_repeats=2)
    assert st["shots"] == 2 * S and st["sweeps"] == 2 * S and st["repeats"] == 2 and not st["steady_state"]
    H0 = np.fft.ifftshift(shot)
    assert st["H"].shape == K + (2,)
    assert np.array_equal(st["H"][..., 0], H0)
    assert np.array_equal(st["H"][..., 1], np.where(H0 > 0, H0 + S, 0))
    T = st["time_index"]
    assert np.array_equal(T[..., 1], np.where(T[..., 0] > 0, T[..., 0] + NP, 0))
    ref = TO.trajectory_states(k2p, kRange, NP // S, NRepeats=2)
    for key in ("H", "time_index", "sweep_of_profile"):
        assert np.array_equal(st[key], ref[key]), key
    k1, k2, off, st2 = TR.sampling_from_trajectory(k2p, kRange, NP // S, n_repeats=2)
    rep = st2["sample_repeat"]
    assert np.array_equal(np.diff(off), np.full(2 * S, NP // S))
    state = np.repeat(np.arange(1, 2 * S + 1), NP // S)
    assert np.array_equal(rep, (state > S).astype(rep.dtype))
    assert (st["H"][k1, k2, rep] == state).all()
    # the same lines in the same order in both passes
    assert np.array_equal(k1[: off[S]], k1[off[S]:]) and np.array_equal(k2[: off[S]], k2[off[S]:])
    # against the oracle's buildHarmonicSampling over the 3-D state map
    from oracle import disorder_oracle as O
    ok1, ok2, ooff, oidx = O.build_harmonic_sampling(st["H"], 2 * S)
    assert np.array_equal(ok1[: ooff[-2]], k1) and np.array_equal(ok2[: ooff[-2]], k2)
    assert np.array_equal(oidx[: ooff[-2]] // (K[0] * K[1]), rep)
    with pytest.raises(ValueError):
        TR.trajectory_states(np.concatenate([ko, ko, ko]), kRange, NP // S, n_repeats=2)
    # one repeat: the second pass has no slot
    with pytest.raises(ValueError):
        TR.trajectory_states(k2p, kRange, NP // S)


@pytest.mark.parametrize("R", [2, 3])
def test_repeats_single_shot_steady_state(R):
    """One echo train and R repeats (solveXT.m:85-90, :110-126): the
    trajectory is repeated R times, the scan is steady state (R shots / R
    repeats = 1) and the DCT of the repeated triangle wave peaks at R times
    the single scan's frequency, so 2PR sweeps of M profiles; repeat r's
    profiles fill slot r."""
    P, M = 2, 12
    c = 2 * P
    up = np.arange(M) - M // 2
    tri = np.concatenate([up, up[::-1]] * P)
    half = np.repeat(np.arange(2 * P), M)
    kTraj = np.stack([tri, c * tri + half], 1)
    kRange = [[-M // 2, M // 2 - 1], [int(kTraj[:, 1].min()), int(kTraj[:, 1].max())]]
    NP = len(tri)
    a = TR.trajectory_states(kTraj, kRange, 1, n_repeats=R)
    b = TO.trajectory_states(kTraj, kRange, 1, NRepeats=R)
    assert a["steady_state"] and b["steady_state"] and a["shots"] == R
    assert a["sweeps"] == b["sweeps"] == 2 * P * R
    assert np.array_equal(a["sweep_of_profile"], np.repeat(np.arange(1, 2 * P * R + 1), M))
    for key in ("H", "time_index", "sweep_of_profile"):
        assert np.array_equal(a[key], b[key]), key
    T = a["time_index"]
    for r in range(1, R):
        assert np.array_equal(T[..., r], np.where(T[..., 0] > 0, T[..., 0] + r * NP, 0))
    assert np.array_equal(a["H"][..., R - 1], np.where(a["H"][..., 0] > 0, a["H"][..., 0] + 2 * P * (R - 1), 0))
    with pytest.raises(ValueError):  # two shots for three repeats
        TR.trajectory_states(kTraj, kRange, NP // 2, n_repeats=3)
