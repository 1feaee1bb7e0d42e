"""Literal FP64 restatement of the simulation-side operators
(Simulation/encodeDISORDER.m, decodeDISORDER.m, errorFit.m,
precondDISORDER.m, Xsolver.m, gradientDISORDER.m, directionDISORDER.m) on
dense per-shot k-space planes with full-plane masks, and of one pass of the
joint loop of Simulation/pyramidDISORDER.m:62-132.

TEST INFRASTRUCTURE ONLY: the checker of disorder_amd.simulator; the
product never imports it.

Arrays: x (n1, n2, n3); S (n1, n2, n3, C); masks A (n1, n2, NS) with DC
first; dense data y (n1, n2, n3, C, NS); T (NS, 6).  Single repeat, no
multiband, identity field-of-view transform.
Pinning (tests/test_simulator.py): decode is the adjoint of encode divided
by n1 n2 (dot test); errorFit is the squared norm of the masked residual;
Xsolver converges to the dense least-squares solution and counts its
operations as the MATLAB loop does; the direction is the damped
Gauss-Newton step of a quadratic whose minimiser is known.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import disorder_oracle as O


def _transform_shots(x, T, Nfull):
    """T_s x for every shot: (n1, n2, n3, NS)."""
    return O.rigid_forward(x, T, x.shape[:3], Nfull)


def encode_sim(x, S, A, T, Nfull, y=None):
    """encodeDISORDER(x, S, et(T), [], A, [], y): T, S, fft over dims 1-2
    (unnormalised), minus y, times A.  Returns (n1, n2, n3, C, NS)."""
    Xs = _transform_shots(x, T, Nfull)                           # (n1,n2,n3,NS)
    v = Xs[:, :, :, None, :] * S[:, :, :, :, None]               # (n1,n2,n3,C,NS)
    v = np.fft.fft(np.fft.fft(v, axis=0), axis=1)
    if y is not None:
        v = v - y
    return v * A[:, :, None, None, :]


def decode_sim(y, S, A, T, Nfull, M=None):
    """decodeDISORDER(y, conj(S), et(T, inverse), [], A, [], M): times A,
    ifft over dims 1-2, times conj(S) summed over coils, inverse transform
    summed over shots, times M."""
    v = y * A[:, :, None, None, :]
    v = np.fft.ifft(np.fft.ifft(v, axis=0), axis=1)
    v = np.sum(v * np.conj(S)[:, :, :, :, None], axis=3)          # (n1,n2,n3,NS)
    x = O.rigid_inverse(v, T, v.shape[:3], Nfull, sumR=True)
    return x if M is None else x * M


def error_fit(x, y, S, A, T, Nfull, per_shot=False):
    """errorFit.m: |encode(x) - y|^2 under the mask, summed (or per shot)."""
    r = np.abs(encode_sim(x, S, A, T, Nfull, y)) ** 2
    return r.sum(axis=(0, 1, 2, 3)) if per_shot else float(r.sum())


def precond(S, reg=1e-3):
    return 1.0 / (np.sum(np.abs(S) ** 2, axis=3) + reg)


def xsolver(x, y, S, A, T, Nfull, M=None, nX=100, toler=1e-8, lam=0.0):
    """Xsolver.m:29-92 (A given to decode, NY(1:2) to encode: the mask acts once)."""
    P = precond(S)

    def system(v):
        return decode_sim(encode_sim(v, S, np.ones_like(A), T, Nfull), S, A, T, Nfull, M) + lam * v
    n = 0
    if x is not None:
        r = decode_sim(y, S, A, T, Nfull, M) - system(x)
        n += 3
    else:
        r = -decode_sim(y, S, A, T, Nfull, M)
        n += 1
        x = np.zeros_like(r)
    z = P * r
    p = z.copy()
    rsold = np.sum(np.conj(z) * r)
    while True:
        Ap = system(p)
        n += 2
        al = np.conj(rsold) / np.sum(np.conj(p) * Ap)
        xup = al * p
        x = x + xup
        if np.max(np.abs(xup) ** 2) < toler or n >= nX:
            break
        r = r - al * Ap
        z = P * r
        rsnew = np.sum(np.conj(z) * r)
        be = rsnew / rsold
        p = z + be * p
        rsold = rsnew
        if math.sqrt(abs(rsnew)) < 1e-10:
            break
        n += 1
    return x, n


HA_SIM = [(1, 1), (2, 2), (3, 3), (1, 2), (1, 3), (2, 3), (1, 4), (2, 4), (3, 4), (1, 5), (2, 5), (3, 5),
          (1, 6), (2, 6), (3, 6), (4, 5), (4, 6), (5, 6), (4, 4), (5, 5), (6, 6)]


def gradient(x, y, S, A, T, Nfull):
    """gradientDISORDER.m with the Jacobian of sincRigidTransformGradient:
    ry = encode(T x) - y, J_g = encode(dT x/dT_g) (no further motion);
    dt [6 x NS] = sum Re(J conj ry), ddt [21 x NS] over HA_SIM."""
    NS = T.shape[0]
    dt = np.zeros((6, NS))
    ddt = np.zeros((21, NS))
    EX = np.zeros(NS)
    n = x.shape[:3]
    for s in range(NS):
        f = O.Factors(T[s:s + 1], n, Nfull, 1, cg=True)
        xt, xB = O.sinc_rigid_transform(x, f, 1, keep=True)
        G = O.sinc_rigid_transform_gradient(xB, f)
        As = A[:, :, s:s + 1]
        Z = np.zeros((1, 6))

        def enc(v):
            return encode_sim(v[..., 0] if v.ndim == 4 else v, S, As, Z, Nfull)
        ry = enc(xt) - y[..., s:s + 1] * As[:, :, None, None, :]
        EX[s] = float(np.sum(np.abs(ry) ** 2))
        J = [enc(G[g]) for g in range(6)]
        for g in range(6):
            dt[g, s] = float(np.sum(np.real(J[g] * np.conj(ry))))
        for k, (a, b) in enumerate(HA_SIM):
            ddt[k, s] = float(np.sum(np.real(J[a - 1] * np.conj(J[b - 1]))))
    return EX, dt, ddt


def direction(dt, ddt, w=None, flagw=None, winic=None):
    """directionDISORDER.m:15-58."""
    maxTDeg = 1 * math.pi / 180
    multA, multB = 1.2, 2.0
    NS = dt.shape[1]
    if w is None:
        w = np.full(NS, 1e-3) if winic is None else np.full(NS, winic)
    else:
        w = np.array(w, dtype=np.float64)
    if flagw is None:
        flagw = np.zeros(NS)
    else:
        w[flagw == 1] = w[flagw == 1] / multA
        w[flagw == 0] = w[flagw == 0] * multB
    Tup = np.zeros((NS, 6))
    dHs = []
    for s in range(NS):
        H = np.eye(6)
        for k, (a, b) in enumerate(HA_SIM):
            if a == b:
                H[a - 1, b - 1] = ddt[k, s] * (1 + w[s])
            else:
                H[a - 1, b - 1] = ddt[k, s]
                H[b - 1, a - 1] = ddt[k, s]
        H = H + np.max(np.abs(H)) / 1e6 * np.eye(6)
        st = np.linalg.solve(H, dt[:, s])
        dHs.append(st / w[s])
    dH = np.stack(dHs, axis=1)
    if winic is None:
        winic = maxTDeg / np.max(np.abs(dH[3:6]))
    dH = winic * dH
    return dH.T.copy(), w, np.zeros(NS), winic


def restrict(T, lim):
    """restrictTransform(T, limTran) (Methods/restrictTransform.m:20-33)."""
    T = np.array(T, dtype=np.float64, copy=True)
    for j in range(3, 6):
        a = T[:, j]
        a[(a >= math.pi) | (a <= -math.pi)] = 0.0
    for d in range(3):
        t = T[:, d]
        t[(t >= lim[d][1]) | (t <= lim[d][0])] = 0.0
    return T


def joint_iteration(x, y, S, A, T, Nfull, state, lim, M=None, meanT=True):
    """pyramidDISORDER.m:77-131 (one pass of the while loop, after its
    convergence test).  Returns (x, T, nops) with nops in units of NT."""
    EX, dt, ddt = gradient(x, y, S, A, T, Nfull)
    nops = 1
    w, flagw, winic = state.get("w"), state.get("flagw"), state.get("winic")
    if w is None or np.any(flagw == 1):
        state["prev"] = (x.copy(), T.copy(), EX.copy(), dt.copy(), ddt.copy())
        nops += 6
    else:
        x, T, EX, dt, ddt = [np.array(v, copy=True) for v in state["prev"]]
    Tup, w, flagw, winic = direction(dt, ddt, w, flagw, winic)
    Tca = restrict(T - Tup, lim)
    ET = error_fit(x, y, S, A, Tca, Nfull, per_shot=True)
    nops += 1
    flagw = np.where(ET <= EX, 1.0, 0.0)
    Tn = np.where(flagw[:, None] == 1, Tca, T)
    ry = encode_sim(x, S, A, Tn, Nfull, y)
    xu, nu = xsolver(None, ry, S, A, Tn, Nfull, M)
    x = x + xu
    nops += nu
    if meanT:
        Tmed = Tn.mean(axis=0, keepdims=True)
        x = O.rigid_forward(x, Tmed, x.shape[:3], Nfull)[..., 0]
        if M is not None:
            x = x * M
        Tn = Tn - Tmed
    state.update({"w": w, "flagw": flagw, "winic": winic})
    return x, Tn, nops
