"""Literal CPU restatement of the DISORDER trajectory code (integer, bit-exact).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
disorder_amd.trajectory; the product never imports it.

Follows, loop for loop and with 1-based indices kept as in the MATLAB text:
  Acquisition/electrostaticRepulsionDISORDER.m:10-47   electrostatic_order
  Acquisition/samplingDISORDER.m:31-198                 sampling_disorder
  Reconstruction/solveXT.m:75-141, :349-350             trajectory_states
  Reconstruction/solveXT.m:270-287                      subdivide_states
Random permutations (randperm) are drawn from a numpy Generator in the
reference's order.  Pinning: hand-worked electrostatic orders for U = 2x2
and 3x3, closed forms of the orders on full grids (one point per shot per
tile, echo counts, zig-zag key order), the DCT frequency of a sweep
trajectory (also repeated R times), and the repeat slots of a scan acquired
twice (slot 2 = slot 1 shifted by one pass); see tests/test_trajectory.py.
"""
from __future__ import annotations

import math

import numpy as np


def _ind2sub(N, ind):
    """MATLAB ind2sub for 2-D, 1-based in and out."""
    i = (ind - 1) % N[0] + 1
    j = (ind - 1) // N[0] + 1
    return i, j


def electrostatic_order(N):
    N = [int(N[0]), int(N[1])]
    P = N[0] * N[1]
    poi = np.zeros((P, 2))
    ind = np.zeros(P, np.int64)
    free = np.ones(P + 1, bool)
    ind[0] = (math.ceil(N[0] / 2)) + (math.ceil(N[1] / 2) - 1) * N[0]
    poi[0] = _ind2sub(N, ind[0])
    free[ind[0]] = False
    for n in range(2, P + 1):
        best, bval = None, math.inf
        for m in range(1, P + 1):
            if not free[m]:
                continue
            c = _ind2sub(N, m)
            wrepul = 0.0
            for q in range(n - 1):
                wd = math.inf
                for s1 in (-N[0], 0, N[0]):
                    for s2 in (-N[1], 0, N[1]):
                        d = (c[0] - (poi[q, 0] + s1)) ** 2 + (c[1] - (poi[q, 1] + s2)) ** 2
                        wd = min(wd, d)
                wrepul += 1.0 / wd
            if wrepul < bval - 1e-9:
                best, bval = m, wrepul
        ind[n - 1] = best
        poi[n - 1] = _ind2sub(N, best)
        free[best] = False
    return ind


def sampling_disorder(k, U=None, SegmentOrder="RandomCheckered", SampleOrder="AlternatingZigZag", K=None, rng=None):
    k = [tuple(int(v) for v in row) for row in np.asarray(k)]
    NP = len(k)
    if U is None:
        U = [round(math.sqrt(NP))] * 2
    U = [int(U[0]), int(U[1])]
    if K is None:
        K = [max(p[d] for p in k) - min(p[d] for p in k) + 1 for d in range(2)]
    K = [int(K[0]), int(K[1])]
    Uor, SegOr = list(U), SegmentOrder
    if SegmentOrder == "Sequential":
        SegmentOrder, U, k, K = "Checkered", [1, 1], [(p[1], p[0]) for p in k], K[::-1]
    S = U[0] * U[1]
    kGrid = [[j - math.ceil((K[d] + 1) / 2) for j in range(1, K[d] + 1)] for d in range(2)]
    Ul = []
    for d in range(2):
        st = [g for g in kGrid[d] if (g - math.ceil(U[d] / 2)) % U[d] == 0]
        st = [st[0] - (st[1] - st[0])] + st
        Ul.append([(a, a + U[d] - 1) for a in st])
    UperK = [len(Ul[0]), len(Ul[1])]
    mk = [min(p[d] for p in k) - 1 for d in range(2)]
    ElSh = set((p[0] - mk[0], p[1] - mk[1]) for p in k)
    NPBlock = np.zeros(UperK, bool)
    for m in range(1, UperK[0] + 1):
        for n in range(1, UperK[1] + 1):
            l1, l2 = Ul[0][m - 1], Ul[1][n - 1]
            if l1[0] >= kGrid[0][0] and l1[1] <= kGrid[0][-1] and l2[0] >= kGrid[1][0] and l2[1] <= kGrid[1][-1]:
                corn = [(l1[0], l2[0]), (l1[1], l2[0]), (l1[0], l2[1]), (l1[1], l2[1])]
                NPBlock[m - 1, n - 1] = sum((c[0] - mk[0], c[1] - mk[1]) in ElSh for c in corn) == 4
    posAcc = []
    for lin in range(1, UperK[0] * UperK[1] + 1):
        m, n = _ind2sub(UperK, lin)
        if NPBlock[m - 1, n - 1]:
            posAcc.append((m, n))
    assign = [0] * NP
    for m, n in posAcc:
        l1, l2 = Ul[0][m - 1], Ul[1][n - 1]
        for i, p in enumerate(k):
            if l1[0] <= p[0] <= l1[1] and l2[0] <= p[1] <= l2[1]:
                assign[i] = 1
    kk = []
    for m, n in posAcc:
        for lin in range(1, S + 1):
            a, b = _ind2sub(U, lin)
            kk.append((a + Ul[0][m - 1][0] - 1, b + Ul[1][n - 1][0] - 1))
    out = [p for i, p in enumerate(k) if not assign[i]]
    keys = []
    for p in out:
        x1, x2 = np.float32(p[0]), np.float32(p[1])
        ang = np.float32(-np.arctan2(x2, x1))
        rad = np.float32(np.sqrt(np.float32(x1 * x1 + x2 * x2)))
        keys.append(np.float32(ang + np.float32(np.float32(1e-5) * rad)))
    for idx in sorted(range(len(out)), key=lambda i: (keys[i], i)):
        kk.append(out[idx])
    Srand = list(electrostatic_order(U))
    for e in range(1, S * (NP // S) + 1, S):
        vE = list(range(e, e + S))
        if SegmentOrder == "RandomCheckered":
            Srand = list(rng.permutation(S) + 1)
        old = [kk[v - 1] for v in vE]
        for j in range(S):
            kk[vE[Srand[j] - 1] - 1] = old[j]
    if SegmentOrder == "Random":
        pr = rng.permutation(NP)
        kk = [kk[i] for i in pr]
    for s in range(1, S + 1):
        vS = list(range(s, NP + 1, S))
        kke = [list(kk[v - 1]) for v in vS]
        ind = []
        for row in kke:
            found = (0, 0)
            ent = False
            for m in range(1, UperK[0] + 1):
                if Ul[0][m - 1][0] <= row[0] <= Ul[0][m - 1][1]:
                    for n in range(1, UperK[1] + 1):
                        if Ul[1][n - 1][0] <= row[1] <= Ul[1][n - 1][1]:
                            found = (m, n)
                            if m % 2 == 0:
                                row[1] = -row[1]
                                found = (m, UperK[1] + 1 - n)
                                ent = True
                                break
                    if ent:
                        break
            ind.append(found)
        or2 = [1e6 * f[1] for f in ind]
        if SegOr == "Sequential" and SampleOrder == "ZigZag":
            pr = rng.permutation(len(or2))
            or2 = [or2[i] for i in pr]
        key = [1e9 * ind[i][0] + or2[i] + 1e3 * kke[i][0] + kke[i][1] for i in range(len(vS))]
        srt = sorted(range(len(vS)), key=lambda i: (key[i], i))
        if s % 2 == 0 and SampleOrder == "AlternatingZigZag":
            srt = srt[::-1]
        if SegmentOrder != "Random":
            old = [kk[v - 1] for v in vS]
            for i, v in enumerate(vS):
                kk[v - 1] = old[srt[i]]
    kout = []
    for s in range(1, S + 1):
        kout += [kk[v - 1] for v in range(s, NP + 1, S)]
    if SegOr == "Sequential":
        kout = [(p[1], p[0]) for p in kout]
        K = K[::-1]
        kGrid = [[j - math.ceil((K[d] + 1) / 2) for j in range(1, K[d] + 1)] for d in range(2)]
        U = Uor
        S = U[0] * U[1]
        if SampleOrder == "ZigZag":
            NE = NP // S
            for s in range(1, S + 1):
                blk = kout[(s - 1) * NE:s * NE]
                key = [1e3 * p[0] + (-p[1] if p[0] % 2 == 0 else p[1]) for p in blk]
                srt = sorted(range(NE), key=lambda i: (key[i], i))
                kout[(s - 1) * NE:s * NE] = [blk[i] for i in srt]
    S = Uor[0] * Uor[1]
    Echo = np.zeros(K, np.int64)
    Shot = np.zeros(K, np.int64)
    cont = 0
    for s in range(1, S + 1):
        for e in range(1, len(range(s, NP + 1, S)) + 1):
            p = kout[cont]
            Echo[p[0] - kGrid[0][0], p[1] - kGrid[1][0]] = e
            Shot[p[0] - kGrid[0][0], p[1] - kGrid[1][0]] = s
            cont += 1
    return np.array(kout, np.int64), NP, np.array(K), kGrid, Shot, Echo


def dct2_ortho(x):
    """Orthonormal DCT-II along axis 0 written from its definition
    (Methods/fct.m computes the same transform through an FFT)."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    j = np.arange(n)
    C = np.cos(np.pi * (2 * j[None, :] + 1) * j[:, None] / (2 * n)) * math.sqrt(2.0 / n)
    C[0] /= math.sqrt(2.0)
    return C @ x


def trajectory_states(kTraj, kRange, NEchos, groupSweeps=1, NRepeats=1):
    kTraj = np.asarray(kTraj, np.int64)
    NProfs = kTraj.shape[0]
    if NEchos == 1:
        NEchos = NProfs
    NShots = NProfs // NEchos
    if NShots < NRepeats:  # solveXT.m:85-90
        NShots = NRepeats
        NProfs = NShots * NEchos
        kTraj = np.concatenate([kTraj] * NRepeats, axis=0)
        NEchos = NProfs // NShots
    kShift = [math.floor((kRange[d][1] - kRange[d][0] + 1) / 2) + 1 for d in range(2)]
    kSize = [kRange[d][1] - kRange[d][0] + 1 for d in range(2)]
    steady = False  # solveXT.m:110-115
    if NShots % NRepeats == 0:
        steady = NShots // NRepeats == 1
    if steady:
        A = np.abs(dct2_ortho(kTraj))
        iAs = [int(np.argmax(A[:NProfs // 2, c])) + 1 for c in range(2)]
        iM = min(iAs)
        NSamples = NProfs / (iM - 1)
        v = NProfs / (NSamples * groupSweeps)
        NSweeps = int(math.floor(v + 0.5))
    else:
        NSweeps = NShots
    sweep = [math.ceil(NSweeps * ((n - 0.5) / NProfs)) for n in range(1, NProfs + 1)]
    T = np.zeros(kSize + [NRepeats], np.int64)  # solveXT.m:132-141
    for n in range(1, kTraj.shape[0] + 1):  # size(kIndex, 1)
        i, j = kTraj[n - 1, 0] + kShift[0], kTraj[n - 1, 1] + kShift[1]
        for s in range(NRepeats):
            if T[i - 1, j - 1, s] == 0:
                T[i - 1, j - 1, s] = n
                break
    for m in range(2):
        T = np.fft.ifftshift(T, axes=m)
    H = np.zeros_like(T)
    for i in range(T.shape[0]):
        for j in range(T.shape[1]):
            for s in range(NRepeats):
                if T[i, j, s]:
                    H[i, j, s] = sweep[T[i, j, s] - 1]
    if NRepeats == 1:
        T, H = T[..., 0], H[..., 0]
    return {"shots": NShots, "echoes": NEchos, "sweeps": NSweeps, "steady_state": steady,
            "sweep_of_profile": np.array(sweep), "time_index": T, "H": H}


def subdivide_states(stateSample, subT):
    st = list(int(v) for v in stateSample)
    c = 0
    for n in range(1, len(subT) + 1):
        if subT[n - 1] == 2:
            ind = [i for i, v in enumerate(st) if v == n + c]
            first = ind[len(ind) // 2]
            for i in range(first, len(st)):
                st[i] += 1
            c += 1
    return np.array(st)
