"""The simulation-side reconstruction operators (Simulation/encodeDISORDER.m,
decodeDISORDER.m, errorFit.m, precondDISORDER.m, Xsolver.m,
gradientDISORDER.m, directionDISORDER.m) and one joint iteration of the
loop of Simulation/pyramidDISORDER.m:62-132 that uses them, on the GPU.

The simulator's model is the reconstruction's model with an unnormalised
DFT along the two phase encodes and full-plane per-shot masks A: for a
volume with n1 x n2 phase encodes, encodeDISORDER = sqrt(n1 n2) E and
decodeDISORDER = E^H / sqrt(n1 n2) (ifft carries the 1/(n1 n2)), E the
unitary encoding of disorder_amd.ops.Encoding.  The masked k-space planes
are kept compact: only the sampled lines, in the state order of the
harmonic sampling, so A is the sampling itself (zeros elsewhere are not
stored).  Single repeat, no multiband (B empty) and the identity
field-of-view transform (TFov = 0 for one repeat, synthesizeT.m:21-28).
"""
from __future__ import annotations

import math

import numpy as np

from . import ops

# second-order derivative pairs of gradientDISORDER.m:23-24 (1-based, as sincRigidTransformHessian orders them)
HA_SIM = [(1, 1), (2, 2), (3, 3), (1, 2), (1, 3), (2, 3), (1, 4), (2, 4), (3, 4), (1, 5), (2, 5), (3, 5),
          (1, 6), (2, 6), (3, 6), (4, 5), (4, 6), (5, 6), (4, 4), (5, 5), (6, 6)]


class SimOperator:
    """Operators of one volume: maps S [n1,n2,n3,C] (device), the shots'
    sampled lines k1, k2 with offsets off, the support mask M (device float
    [n1,n2,n3] or None), Nfull the grid of the motion parameters."""

    def __init__(self, S, k1, k2, off, Nfull=None, M=None, reg=1e-3, ctx=None):
        self.E = ops.Encoding(S, k1, k2, off, Nfull=Nfull, ctx=ctx)
        self.S = S
        self.n = tuple(int(v) for v in S.shape[:3])
        self.s = math.sqrt(self.n[0] * self.n[1])
        self.M = M
        self.P = precond(S, reg, self.E.ctx)
        self.off = np.asarray(off)

    def close(self):
        self.E.close()

    def encode(self, x, T, y=None):
        """encodeDISORDER(x, S, et(T), [], A, [], y): sqrt(n1 n2) E x (- y)."""
        out = self.E.encode(x, T)
        out.mul_(self.s)
        if y is not None:
            out.sub_(y)
        return out

    def decode(self, y, T):
        """decodeDISORDER(y, conj(S), et(T, inverse), [], A, [], M)."""
        x = self.E.decode(y, T)
        x.mul_(1.0 / self.s)
        if self.M is not None:
            x.mul_(self.M)
        return x

    def error_fit(self, x, y, T, per_state=False):
        """errorFit.m: sum |A (F S T x - y)|^2 (total, or per shot with sep 5)."""
        ys = y * (1.0 / self.s)
        ps = self.E.energy(x, ys, T, per_sample=True) * (self.s ** 2)
        if per_state:
            return np.add.reduceat(ps, self.off[:-1]) if len(ps) else np.zeros(len(self.off) - 1)
        return float(np.sum(ps))

    def xsolver(self, y, T, x=None, nX=100, toler=1e-8, lam=0.0):
        """Xsolver.m: CG on decode(encode(x)) + lam x = decode(y) (x given) or
        = -decode(y) from zero (x None), preconditioner 1/(sum|S|^2 + 1e-3),
        stop when max|alpha p|^2 < toler, sqrt|rs| < 1e-10 or when the
        operation count n (3 or 1 to start, 3 per iteration, the last one 2)
        reaches nX.  Returns (x, n)."""
        n0 = 3 if x is not None else 1
        kmax = max(1, math.ceil((nX - n0 + 1) / 3))
        yu = y * ((1.0 if x is not None else -1.0) / self.s)
        xo, _, it = ops.CGsolver(yu, self.E, T, x=x, nIt=kmax, tolType="MaxUpdate", tol=toler, Se=self.P, Ma=self.M,
                                 lam=lam)
        return xo, (n0 + 3 * it - 1 if it > 0 else n0)

    def partials(self, x, y, T):
        """(EX [S], dt [6 x S], ddt [21 x S]) of gradientDISORDER.m with the
        residual ry = encode(T x) - y: energies per shot, sum Re(J conj ry),
        sum Re(J_a conj J_b) over the HA_SIM pairs (unnormalised scale)."""
        En, dG, dH = ops.lm_partials(y * (1.0 / self.s), self.E, x, T)
        nel = self.E.Ns * self.E.n[2] * self.E.C
        s2 = self.s ** 2
        from itertools import combinations
        ha_lm = [(g, g) for g in range(6)] + list(combinations(range(6), 2))
        idx = {p: i for i, p in enumerate(ha_lm)}
        ddt = np.stack([dH[:, idx[tuple(sorted((a - 1, b - 1)))]] for a, b in HA_SIM]) * s2
        return En * nel * s2, dG.T * s2, ddt


def precond(S, reg=1e-3, ctx=None):
    """precondDISORDER.m: P = 1 / (sum_c |S_c|^2 + reg)."""
    return ops.coil_precond(S, reg, ctx)


def direction(dt, ddt, w=None, flagw=None, winic=None, maxTDeg=math.pi / 180, multA=1.2, multB=2.0):
    """directionDISORDER.m: per shot the damped Gauss-Newton step in FP64.
    First call w = winic (or 1e-3); then w /= 1.2 where the energy fell and
    *= 2 where it did not; diagonal times (1 + w), plus max|H|/1e6 I;
    step = H \\ dt / w, scaled once by winic = 1 deg / max|rotation step|.
    Returns (Tup [S x 6], w, flagw (zeroed), winic)."""
    dt = np.asarray(dt, np.float64)
    ddt = np.asarray(ddt, np.float64)
    S = dt.shape[1]
    if w is None:
        w = np.full(S, 1e-3 if winic is None else winic, np.float64)
    else:
        w = np.array(w, np.float64, copy=True)
    if flagw is None:
        flagw = np.zeros(S)
    else:
        flagw = np.asarray(flagw)
        w = np.where(flagw == 1, w / multA, w * multB)
    H = np.zeros((S, 6, 6))
    for k, (a, b) in enumerate(HA_SIM):
        if a == b:
            H[:, a - 1, a - 1] = ddt[k] * (1 + w)
        else:
            H[:, a - 1, b - 1] = ddt[k]
            H[:, b - 1, a - 1] = ddt[k]
    # identity where no pair is set (ddtd starts as eye), then the stabiliser
    eye = np.eye(6)
    set_ = np.zeros((6, 6), bool)
    for a, b in HA_SIM:
        set_[a - 1, b - 1] = set_[b - 1, a - 1] = True
    H = np.where(set_[None], H, eye[None])
    H = H + (np.abs(H).max(axis=(1, 2)) / 1e6)[:, None, None] * eye[None]
    st = np.stack([np.linalg.solve(H[s], dt[:, s]) for s in range(S)], axis=1)
    dH = st / w[None, :]
    if winic is None:
        winic = maxTDeg / np.abs(dH[3:6]).max()
    dH = winic * dH
    return dH.T.copy(), w, np.zeros(S), winic


def restrict(T, lim_tran):
    """restrictTransform(T, limTran): rotations outside (-pi, pi) and
    translations outside (lo, hi) set to zero (translation 3 with [0 0]
    is always zero)."""
    T = np.array(T, np.float64, copy=True)
    ang = T[:, 3:6]
    ang[(ang >= math.pi) | (ang <= -math.pi)] = 0.0
    for d in range(3):
        lo, hi = lim_tran[d]
        t = T[:, d]
        t[(t >= hi) | (t <= lo)] = 0.0
    return T


def joint_iteration(op: SimOperator, x, y, T, state, lim_tran, meanT=True):
    """One pass of the while loop of pyramidDISORDER.m:62-132 at one level:
    residual and energies EX at T (:81-84), gradients when the last step
    was accepted or on the first pass (:85-92, else the previous state is
    restored), direction (:99-102), trial motion restricted (:103-106),
    trial energies ET and per-shot acceptance (:107-112), the image update
    Xsolver([], ry) (:121-122) and the mean-motion removal (:123-131).
    state: dict carried between calls (w, flagw, winic, prev); returns
    (x, T, nops) with nops the effective-operation count of the pass in
    units of NT transforms (1 + [6] + 1 + Xsolver's n)."""
    import torch
    S = T.shape[0]
    EX, dt, ddt = op.partials(x, y, T)
    nops = 1
    w, flagw, winic = state.get("w"), state.get("flagw"), state.get("winic")
    if w is None or (flagw == 1).any():
        state["prev"] = (x.clone(), T.copy(), EX.copy(), dt.copy(), ddt.copy())
        nops += 6
    else:
        xp, Tp, EXp, dtp, ddtp = state["prev"]
        x, T, EX, dt, ddt = xp.clone(), Tp.copy(), EXp, dtp, ddtp
    Tup, w, flagw, winic = direction(dt, ddt, w, flagw, winic)
    Tca = restrict(T - Tup, lim_tran)
    ET = op.error_fit(x, y, Tca, per_state=True)
    nops += 1
    flagw = (ET <= EX).astype(float)
    Tn = np.where(flagw[:, None] == 1, Tca, T)
    # residual of the accepted motion per shot (ry at T for the rejected shots)
    ry = op.encode(x, Tn, y)
    xu, nu = op.xsolver(ry, Tn, None)
    x = x + xu
    nops += nu
    if meanT:
        Tmed = Tn.mean(axis=0, keepdims=True)
        x = ops.sincRigidTransform(x, Tmed, 1, Nfull=op.E.Nfull, ctx=op.E.ctx)[..., 0]
        x = ops.colmajor(x.contiguous()) if not ops.is_colmajor(x) else x
        if op.M is not None:
            x = x * op.M
        Tn = Tn - Tmed
    state.update({"w": w, "flagw": flagw, "winic": winic})
    torch.cuda.synchronize()
    return x, Tn, nops
