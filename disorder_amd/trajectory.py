"""Scanner trajectories: DISORDER sample orders and the mapping from an
acquired profile list to motion states (host side, integer, bit-exact).

* ``electrostatic_order`` -- Acquisition/electrostaticRepulsionDISORDER.m:10-47:
  a greedy order of the U1 x U2 tile positions, starting at the centre
  (ceil(U/2)), each next position the free one of least summed inverse
  squared distance to the placed ones on the torus of period U (ties within
  1e-9 keep the lowest MATLAB linear index).
* ``sampling_disorder`` -- Acquisition/samplingDISORDER.m:31-198: reorders a
  list of phase-encode points into shots and echoes.  Tiles of U1 x U2
  points wholly inside the elliptical shutter give one point to every shot
  (tile points in column-major order, placed by the electrostatic order, or
  a fresh random permutation per tile for 'RandomCheckered'); the other
  points follow in angular order (key -atan2(k2, k1) + 1e-5 |k|, single
  precision) and take part in the same per-block placement; shot s takes
  every S-th point from s, and its echoes are sorted by (tile row, tile
  column mirrored on even tile rows, k1, k2 negated on even tile rows) --
  the keys 1e9 m + 1e6 n + 1e3 k1 + k2 of :144-149 -- reversed on even
  shots for 'AlternatingZigZag'.  Random draws come from a numpy Generator
  in the order the reference draws them, so a seeded Generator gives a
  reproducible order (the reference draws randperm without a seed).
* ``trajectory_states`` -- Reconstruction/solveXT.m:75-141 and :349-350: the
  profile list in acquisition order gives the number of shots (profiles /
  echoes), the sweeps (shots, or for a steady-state single-shot scan the
  sweep count from the dominant orthonormal DCT-II frequency of the
  trajectory), the sweep of every profile ceil(NSweeps (n - 0.5) / NProfs),
  the time index of every k-space position (its profile, 0 when unsampled,
  DC first after ifftshift) and the state label of every position -- the
  mask that Methods/buildHarmonicSampling.m groups
  (disorder_amd.sampling.harmonic_sampling).
* ``subdivide_states`` -- solveXT.m:270-287: states flagged for a
  within-shot split are cut at their middle profile in acquisition order;
  later states shift up by one.
"""
from __future__ import annotations

import math

import numpy as np


class ReconstructionFail(RuntimeError):
    """A scan the aligned reconstruction declines, with the reason the
    reference prints when it sets rec.Fail = 1 and returns
    (Reconstruction/solveXT.m:33, :77, :83)."""


class InterruptedScan(ReconstructionFail, ValueError):
    """The profiles are not a whole number of shots (solveXT.m:83)."""


class SeparableTrajectory(ReconstructionFail, ValueError):
    """rec.Assign.z{2} and z{3} of different lengths, or a single profile
    (solveXT.m:77)."""


def _mlgrid(K):
    """generateGrid(K, 0, K, ceil((K+1)/2)) per dimension: 1-c .. K-c."""
    return [np.arange(1, k + 1, dtype=np.int64) - (k + 2) // 2 for k in K]


def electrostatic_order(U) -> np.ndarray:
    """1-based MATLAB linear indices into a U1 x U2 grid (column-major)."""
    U1, U2 = int(U[0]), int(U[1])
    n = U1 * U2
    sub = np.stack(np.unravel_index(np.arange(n), (U1, U2), order="F"), axis=1) + 1  # [n, 2] 1-based
    free = np.ones(n, bool)
    first = (math.ceil(U1 / 2) - 1) + (math.ceil(U2 / 2) - 1) * U1
    order = [first]
    free[first] = False
    sh1 = np.array([-U1, 0, U1])
    sh2 = np.array([-U2, 0, U2])

    def add(p):
        d1 = (sub[:, 0:1] - (sub[p, 0] + sh1)[None, :]) ** 2            # [n, 3]
        d2 = (sub[:, 1:2] - (sub[p, 1] + sh2)[None, :]) ** 2
        w = (d1[:, :, None] + d2[:, None, :]).reshape(n, 9).min(axis=1).astype(np.float64)
        with np.errstate(divide="ignore"):
            return 1.0 / w

    rep = add(first)  # summed in placement order, as the reference's sum over placed points
    for _ in range(1, n):
        best, bval = -1, math.inf
        for m in np.flatnonzero(free):
            if rep[m] < bval - 1e-9:
                best, bval = int(m), rep[m]
        order.append(best)
        free[best] = False
        rep = rep + add(best)
    return np.array(order, np.int64) + 1


def _tiles(kg, u):
    """Tile starts/ends along one dimension (samplingDISORDER.m:52-58)."""
    st = kg[np.mod(kg - math.ceil(u / 2), u) == 0]
    if len(st) < 2:
        raise ValueError("the grid must hold at least two tile starts per dimension")
    st = np.concatenate([[st[0] - (st[1] - st[0])], st])
    return np.stack([st, st + u - 1])  # [2, ntiles]


def _tile_of(v, Ul):
    """Index (1-based) of the tile holding each value of v."""
    idx = np.zeros(len(v), np.int64)
    for m in range(Ul.shape[1]):
        hit = (v >= Ul[0, m]) & (v <= Ul[1, m]) & (idx == 0)
        idx[hit] = m + 1
    return idx


SEGMENT_ORDERS = ("Sequential", "Checkered", "RandomCheckered", "Random")
SAMPLE_ORDERS = ("ZigZag", "AlternatingZigZag")


def sampling_disorder(k, U=None, segment_order="RandomCheckered", sample_order="AlternatingZigZag", K=None,
                      rng: np.random.Generator | None = None):
    """(k_out [NP x 2], NP, K, kGrid, ShotNumber [K1 x K2], EchoNumber [K1 x K2])
    for phase-encode points k [NP x 2] (integer grid coordinates)."""
    k = np.asarray(k, np.int64)
    if k.ndim != 2 or k.shape[1] != 2:
        raise ValueError("k must be [NP x 2]")
    if segment_order not in SEGMENT_ORDERS or sample_order not in SAMPLE_ORDERS:
        raise ValueError("unknown segment or sample order")
    NP = k.shape[0]
    U = np.array([round(math.sqrt(NP))] * 2 if U is None else U, np.int64)
    K = (k.max(axis=0) - k.min(axis=0) + 1) if K is None else np.asarray(K, np.int64)
    needs_rng = segment_order in ("RandomCheckered", "Random") or (segment_order == "Sequential"
                                                                     and sample_order == "ZigZag")
    if needs_rng and rng is None:
        raise ValueError(f"{segment_order}/{sample_order} draws random permutations: pass rng")
    Uor, seq = U.copy(), segment_order == "Sequential"
    seg = segment_order
    if seq:
        seg, U, k, K = "Checkered", np.array([1, 1]), k[:, ::-1].copy(), K[::-1].copy()
    S = int(U.prod())
    if sample_order != "AlternatingZigZag" and NP % S:
        raise ValueError(f"number of profiles ({NP}) not a multiple of number of shots ({S})")
    kGrid = _mlgrid(K)
    Ul = [_tiles(kGrid[d], int(U[d])) for d in range(2)]
    UperK = [Ul[0].shape[1], Ul[1].shape[1]]
    # tiles whose four corners are sampled (:61-78)
    mk = k.min(axis=0) - 1
    ElSh = np.zeros((int(K[0]), int(K[1])), bool)
    ElSh[k[:, 0] - mk[0] - 1, k[:, 1] - mk[1] - 1] = True
    inside = np.zeros(UperK, bool)
    for m in range(UperK[0]):
        for n in range(UperK[1]):
            a0, a1 = Ul[0][:, m]
            b0, b1 = Ul[1][:, n]
            if a0 >= kGrid[0][0] and a1 <= kGrid[0][-1] and b0 >= kGrid[1][0] and b1 <= kGrid[1][-1]:
                c = ElSh[np.array([a0, a1, a0, a1]) - mk[0] - 1, np.array([b0, b0, b1, b1]) - mk[1] - 1]
                inside[m, n] = c.all()
    acc = np.argwhere(inside.T)[:, ::-1]  # accepted tiles, column-major tile order
    t1 = _tile_of(k[:, 0], Ul[0])
    t2 = _tile_of(k[:, 1], Ul[1])
    if (t1 == 0).any() or (t2 == 0).any():
        raise ValueError("points outside the spectral grid K")
    assigned = inside[t1 - 1, t2 - 1]
    # tile points, column-major inside each tile, then the boundary points by angle (:81-97)
    tu = np.stack(np.unravel_index(np.arange(S), tuple(U), order="F"), axis=1)
    parts = [tu + np.array([Ul[0][0, m], Ul[1][0, n]]) for m, n in acc]
    out = k[~assigned].astype(np.float32)
    ang = -np.arctan2(out[:, 1], out[:, 0])
    rad = np.sqrt(np.sum(out * out, axis=1, dtype=np.float32))
    key = ang + np.float32(1e-5) * rad
    parts.append(k[~assigned][np.argsort(key, kind="stable")])
    kk = np.concatenate(parts, axis=0) if parts else np.zeros((0, 2), np.int64)
    # checker placement inside blocks of S (:100-106)
    Srand = electrostatic_order(U) - 1
    for e in range(0, S * (NP // S), S):
        if seg == "RandomCheckered":
            Srand = rng.permutation(S)
        blk = kk[e:e + S].copy()
        kk[e + Srand] = blk
    if seg == "Random":
        kk = kk[rng.permutation(NP)]
    # echo orders (:109-152)
    for s in range(S):
        vS = np.arange(s, NP, S)
        kke = kk[vS].copy()
        m = _tile_of(kke[:, 0], Ul[0])
        n = _tile_of(kke[:, 1], Ul[1])
        even = (m % 2) == 0
        k2 = np.where(even, -kke[:, 1], kke[:, 1])
        n = np.where(even, UperK[1] + 1 - n, n)
        or2 = 1e6 * n
        if seq and sample_order == "ZigZag":
            or2 = or2[rng.permutation(len(or2))]
        key = 1e9 * m + or2 + 1e3 * kke[:, 0] + k2
        o = np.argsort(key, kind="stable")
        if s % 2 == 1 and sample_order == "AlternatingZigZag":
            o = o[::-1]
        if seg != "Random":
            kk[vS] = kk[vS[o]]
    kout = np.concatenate([kk[np.arange(s, NP, S)] for s in range(S)], axis=0)
    if seq:
        kout, K = kout[:, ::-1].copy(), K[::-1].copy()
        kGrid = _mlgrid(K)
        U = Uor
        S = int(U.prod())
        if sample_order == "ZigZag":
            NE = NP // S
            for s in range(S):
                blk = kout[s * NE:(s + 1) * NE]
                k2 = np.where(blk[:, 0] % 2 == 0, -blk[:, 1], blk[:, 1])
                kout[s * NE:(s + 1) * NE] = blk[np.argsort(1e3 * blk[:, 0] + k2, kind="stable")]
    S = int(Uor.prod())
    Shot = np.zeros(tuple(int(v) for v in K), np.int64)
    Echo = np.zeros_like(Shot)
    cont = 0
    for s in range(S):
        ne = len(range(s, NP, S))
        pos = kout[cont:cont + ne] - np.array([kGrid[0][0], kGrid[1][0]])
        Echo[pos[:, 0], pos[:, 1]] = np.arange(1, ne + 1)
        Shot[pos[:, 0], pos[:, 1]] = s + 1
        cont += ne
    return kout, NP, K, kGrid, Shot, Echo


def _dct2_ortho(x):
    """Orthonormal DCT-II along axis 0 (Methods/fct.m)."""
    from scipy.fft import dct
    return dct(np.asarray(x, np.float64), type=2, norm="ortho", axis=0)


def trajectory_array(kTraj):
    """[NProfs x 2] phase-encode coordinates from kTraj given as such an
    array or as the pair (rec.Assign.z{2}, rec.Assign.z{3}); a pair of
    different lengths, or a single profile, is a separable y-z trajectory,
    which the aligned reconstruction does not handle (solveXT.m:77)."""
    if isinstance(kTraj, (tuple, list)) and len(kTraj) == 2 and np.ndim(kTraj[0]) == 1:
        z2, z3 = np.asarray(kTraj[0]).ravel(), np.asarray(kTraj[1]).ravel()
        if len(z2) <= 1 or len(z2) != len(z3):
            raise SeparableTrajectory(f"SEPARABLE Y-Z TRAJECTORIES ({len(z2)} and {len(z3)} profiles). "
                                      "ALIGNED RECONSTRUCTION IS NOT PERFORMED")
        kTraj = np.stack([z2, z3], axis=1)
    kTraj = np.asarray(kTraj, np.int64)
    if kTraj.ndim != 2 or kTraj.shape[1] != 2:
        raise ValueError("kTraj must be [NProfs x 2] or the pair (z2, z3)")
    if kTraj.shape[0] <= 1:
        raise SeparableTrajectory("SEPARABLE Y-Z TRAJECTORIES (a single profile). ALIGNED RECONSTRUCTION IS NOT "
                                  "PERFORMED")
    return kTraj


def trajectory_states(kTraj, kRange, n_echoes, group_sweeps=1, n_repeats=1):
    """solveXT.m:75-141, :349-350.

    kTraj [NProfs x 2] phase-encode coordinates in acquisition order; kRange
    [[lo1, hi1], [lo2, hi2]] the k-space extent; n_echoes the echo train
    length (1: the whole scan is one train); n_repeats the averages of the
    data (NY(5)): with fewer shots than repeats the trajectory is repeated
    (:85-90) and each profile takes the first repeat slot of its position
    still free (:132-140).  Returns a dict with shots, echoes, sweeps,
    steady_state, repeats, sweep_of_profile [NProfs] (1-based), time_index
    and H, [NY1 x NY2] for one repeat or [NY1 x NY2 x R] (1-based profile /
    state 1..NSweeps of each position, 0 unsampled, DC first)."""
    kTraj = trajectory_array(kTraj)
    NProfs = kTraj.shape[0]
    NEchos = NProfs if n_echoes == 1 else int(n_echoes)
    if NProfs % NEchos:
        raise InterruptedScan(f"NUMBER OF SHOTS: {NProfs / NEchos:.2f}, PROBABLY AN INTERRUPTED SCAN "
                              f"({NProfs} profiles, {NEchos} echoes per shot)")
    NShots = NProfs // NEchos
    R = int(n_repeats)
    if R < 1:
        raise ValueError("n_repeats must be positive")
    if NShots < R:  # :85-90
        if NShots > 1:  # the repeated list would hold R x NProfs profiles for R x NEchos sweep labels
            raise ValueError(f"{NShots} shots for {R} repeats: fewer shots than repeats needs a single-shot scan")
        NShots = R
        NProfs = NShots * NEchos
        kTraj = np.tile(kTraj, (R, 1))
        NEchos = NProfs // NShots
    kRange = np.asarray(kRange, np.int64)
    kShift = (kRange[:, 1] - kRange[:, 0] + 1) // 2 + 1
    kSize = kRange[:, 1] - kRange[:, 0] + 1
    kIndex = kTraj + kShift[None, :]  # 1-based
    if (kIndex < 1).any() or (kIndex > kSize[None, :]).any():
        raise ValueError("trajectory outside kRange")
    steady = NShots == R  # :112-116, one shot per repeat
    if steady:
        A = np.abs(_dct2_ortho(kTraj))
        iAs = np.argmax(A[:NProfs // 2], axis=0) + 1
        iM = int(iAs.min())
        if iM <= 1:
            raise ValueError("no sweep frequency in a steady-state trajectory")
        NSamples = NProfs / (iM - 1)
        NSweeps = int(np.floor(NProfs / (NSamples * group_sweeps) + 0.5))
    else:
        NSweeps = NShots
    sweep = np.ceil(NSweeps * ((np.arange(1, NProfs + 1) - 0.5) / NProfs)).astype(np.int64)
    T = np.zeros(tuple(int(v) for v in kSize) + (R,), np.int64)
    for n in range(NProfs):
        i, j = kIndex[n] - 1
        free = np.flatnonzero(T[i, j] == 0)
        if len(free) == 0:
            raise ValueError(f"a k-space position is acquired more than {R} time(s)")
        T[i, j, free[0]] = n + 1
    T = np.fft.ifftshift(T, axes=(0, 1))
    H = np.where(T > 0, sweep[np.maximum(T, 1) - 1], 0)
    if R == 1:
        T, H = T[..., 0], H[..., 0]
    return {"shots": NShots, "echoes": NEchos, "sweeps": NSweeps, "steady_state": steady, "repeats": R,
            "sweep_of_profile": sweep, "time_index": T, "H": H}


def subdivide_states(state_of_profile, split):
    """solveXT.m:270-287: state_of_profile [NProfs] (1-based, contiguous
    non-decreasing blocks in acquisition order); split [NStates] bool.
    Each flagged state is cut at its middle profile (the second half starts
    at profile floor(Ni / 2) + 1 of the state) and later states move up by
    one.  Returns (new state_of_profile, new_of_old [NStates] first new
    state of each old one)."""
    st = np.array(state_of_profile, np.int64, copy=True)
    split = np.asarray(split, bool)
    new_of_old = np.zeros(len(split), np.int64)
    c = 0
    for n in range(len(split)):
        new_of_old[n] = n + 1 + c
        if split[n]:
            idx = np.flatnonzero(st == n + 1 + c)
            if len(idx) == 0:
                raise ValueError(f"state {n + 1} has no profiles")
            st[idx[len(idx) // 2]:] += 1
            c += 1
    return st, new_of_old


def sampling_from_trajectory(kTraj, kRange, n_echoes, group_sweeps=1, n_repeats=1):
    """The harmonic sampling of an acquired trajectory: trajectory_states,
    then Methods/buildHarmonicSampling.m on its state map.  Returns (k1, k2,
    off, states) with k1/k2 the 0-based FFT-order indices of the sampled
    lines grouped by state and states the dict of trajectory_states, whose
    'sample_repeat' gives the repeat (0-based) of every sample."""
    from .sampling import harmonic_sampling_repeats
    st = trajectory_states(kTraj, kRange, n_echoes, group_sweeps, n_repeats)
    k1, k2, rep, off = harmonic_sampling_repeats(st["H"], st["sweeps"])
    st["sample_repeat"] = rep
    return k1, k2, off, st
