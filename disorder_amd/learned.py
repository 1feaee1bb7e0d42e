"""Learned motion proposal for the joint reconstruction.

A small network proposes per-shot motion updates from the shots'
sincRigidTransformGradient-based features at the current estimate; the
proposal of a shot is kept only where computeEnergy of that shot decreases,
and LMsolver (Reconstruction/LMsolver.m:1) runs as before on every shot whose
proposal was rejected, so the plain method stays the fallback and decides
what is kept.

Features per shot (13): from the LM partial sums at the current (x, T)
(Reconstruction/LMsolver.m:82-115, via dis_lm_partials):
  [0:6]   diagonal Newton step -g_i / H_ii (translations in voxels,
          rotations in degrees, clipped to +-NEWTON_CLIP),
  [6:12]  log(H_ii / mean over shots of H_ii),
  [12]    log(residual energy / mean over shots).
Proposal: the diagonal Newton step plus the network's output, less its mean
over shots.  Training target of the network: the relative update
(T_true - T_current) minus its mean over shots (the motion is determined
only up to a rigid motion shared by every shot, and the image at the
current estimate sits in the frame of the shots' mean error, so a shot's
energy with that image falls towards the relative update, not the absolute
one), less the Newton step of the features: the network learns the
correction to the diagonal Newton step.

Network, training and inference run in libdisorder_hip.so
(dis_mlp_forward / dis_mlp_backward / dis_mlp_train_step): MLP
13 -> 256 -> 256 -> 6 (GELU, BF16 MFMA products, FP32 weights), then a 1D
convolution along shot order (kernel 5, 6 channels); Adam.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

from . import _lib, ops, synth
from ._lib import check

NIN, NHID, NOUT, NK = 13, 256, 6, 5
DEG = 180.0 / math.pi
NEWTON_CLIP = 20.0  # voxels / degrees
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "motion_net.npy")


def _fp(t):
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
        raise TypeError("contiguous float32 device tensors are required")
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), _lib.c_float_p)


def _offs(scan_off):
    so = np.ascontiguousarray(scan_off, dtype=np.int32)
    return so, so.ctypes.data_as(_lib.c_int32_p), len(so) - 1


class MotionNet:
    """Parameters and Adam state on one device."""

    def __init__(self, params=None, seed=0, ctx: ops.Context = None, device=0):
        self.ctx = ctx or ops.default_context(device)
        self.lib = self.ctx.lib
        self.n = self.lib.dis_mlp_param_count()
        dev = self.ctx.device
        if params is None:
            params = self.init_params(seed)
        p = np.asarray(params, np.float32).ravel()
        if p.size != self.n:
            raise ValueError(f"expected {self.n} parameters, got {p.size}")
        self.p = torch.from_numpy(p.copy()).to(dev)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.step = 0

    @staticmethod
    def init_params(seed):
        """Uniform(+-1/sqrt(fan_in)) weights, zero biases; the convolution
        starts as the identity on its centre tap."""
        rng = np.random.default_rng(seed)
        parts = []
        for fo, fi in ((NHID, NIN), (NHID, NHID), (NOUT, NHID)):
            b = 1.0 / math.sqrt(fi)
            parts += [rng.uniform(-b, b, (fo, fi)).ravel(), np.zeros(fo)]
        Wc = rng.normal(0.0, 0.01, (NOUT, NOUT, NK))
        Wc[np.arange(NOUT), np.arange(NOUT), NK // 2] += 1.0
        parts += [Wc.ravel(), np.zeros(NOUT)]
        return np.concatenate(parts).astype(np.float32)

    @classmethod
    def load(cls, path=WEIGHTS, ctx=None, device=0):
        p = np.load(path, allow_pickle=False)
        return cls(p, ctx=ctx, device=device)

    def save(self, path=WEIGHTS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.p.cpu().numpy())

    def forward(self, X, scan_off):
        X = X.contiguous()
        M = int(X.shape[0])
        Y = torch.empty((M, NOUT), dtype=torch.float32, device=X.device)
        so, sp, nb = _offs(scan_off)
        self.ctx._prep()
        check(self.lib.dis_mlp_forward(self.ctx.h, _fp(self.p), _fp(X), M, sp, nb, _fp(Y)), self.ctx.h)
        return Y

    def backward(self, X, scan_off, dY):
        X = X.contiguous()
        dY = dY.contiguous()
        M = int(X.shape[0])
        G = torch.empty_like(self.p)
        dX = torch.empty_like(X)
        so, sp, nb = _offs(scan_off)
        self.ctx._prep()
        check(self.lib.dis_mlp_backward(self.ctx.h, _fp(self.p), _fp(X), M, sp, nb, _fp(dY), _fp(G), _fp(dX)),
              self.ctx.h)
        return G, dX

    def train_step(self, X, target, scan_off, lr=1e-3):
        self.step += 1
        loss = ctypes.c_float(0.0)
        so, sp, nb = _offs(scan_off)
        self.ctx._prep()
        check(self.lib.dis_mlp_train_step(self.ctx.h, _fp(self.p), _fp(self.m), _fp(self.v), self.step, float(lr),
                                          _fp(X.contiguous()), _fp(target.contiguous()), int(X.shape[0]), sp, nb,
                                          ctypes.byref(loss)), self.ctx.h)
        return loss.value


def features(sums):
    """[S x 28] LM partial sums (energy, 6 gradient, 21 Hessian entries with
    the diagonal first) -> [S x 13] float32 features."""
    s = np.asarray(sums, np.float64)
    en, g, hd = s[:, 0], s[:, 1:7], s[:, 7:13]
    tiny = 1e-30
    hd = np.maximum(hd, tiny)
    F = np.zeros((s.shape[0], NIN))
    step = -g / hd
    step[:, 3:] *= DEG
    F[:, 0:6] = np.clip(np.nan_to_num(step, nan=0.0), -NEWTON_CLIP, NEWTON_CLIP)
    F[:, 6:12] = np.log(hd / np.maximum(hd.mean(axis=0, keepdims=True), tiny))
    F[:, 12] = np.log(np.maximum(en, tiny) / max(en.mean(), tiny))
    return np.nan_to_num(F, nan=0.0, posinf=0.0, neginf=0.0).astype(np.float32)


def to_update(Y):
    """Update in (vox, deg) -> motion update (vox, rad)."""
    d = np.asarray(Y, np.float64).copy()
    d[:, 3:] /= DEG
    return d


def relative_target(T_true, Tc):
    """Training target before the Newton step is taken off: (T_true - Tc)
    less its mean over shots, translations in voxels, rotations in
    degrees."""
    t = np.asarray(T_true, np.float64) - np.asarray(Tc, np.float64)
    t -= t.mean(axis=0, keepdims=True)
    t[:, 3:] *= DEG
    return t.astype(np.float32)


def net_target(F, t):
    """What the network is trained to output for features F and relative
    target t: the correction to the diagonal Newton step."""
    return (np.asarray(t, np.float32) - np.asarray(F, np.float32)[:, 0:6]).astype(np.float32)


def propose(net, F, dev):
    """Proposed update (vox, deg) for one scan's features [S x 13]: the
    Newton step plus the network's correction, less its mean over shots."""
    Y = net.forward(torch.from_numpy(np.ascontiguousarray(F, np.float32)).to(dev), [0, F.shape[0]]).cpu().numpy()
    u = np.asarray(F, np.float64)[:, 0:6] + Y
    return u - u.mean(axis=0, keepdims=True)


def lm_sums(y, E, x, T):
    """Raw per-state LM partial sums [S x 28] (dis_lm_partials)."""
    Tf, Tp = E._T(T)
    sums = np.zeros((E.nmotion, 28), np.float64)
    E.ctx._prep()
    check(E.ctx.lib.dis_lm_partials(E.h, Tp, ops._ptr(x), ops._ptr(y), sums.ctypes.data_as(_lib.c_double_p)),
          E.ctx.h)
    return sums


def state_energy(E, x, y, T, off, Mbe=None):
    ps = E.energy(x, y, T, Mbe=Mbe, per_sample=True)
    return np.add.reduceat(np.asarray(ps, np.float64), off[:-1]) if len(ps) else np.zeros(len(off) - 1)


def level_scale(L):
    """Full-resolution voxels per level voxel (translations of the features
    and updates are expressed in level voxels)."""
    return max(Nf / n for Nf, n in zip(L.E.Nfull, L.n))


class Proposer:
    """Per-iteration proposal step for pyramid.solveXT: returns the motion
    after the proposal and the mask of states that took it (their energy
    decreased); LMsolver then handles the other states.  The network's update
    is tried at the step fractions `alphas` in turn (one energy evaluation
    each, a shot keeps the first fraction that lowers its energy), on the
    first `max_outer` outer iterations of every level only (where the motion
    errors the network was trained on are largest).  With `retire`, a round
    in which no shot keeps its proposal ends the proposals for the rest of
    the solve, so a network that does not help costs one round of features
    and energies, not one per level and outer iteration."""

    def __init__(self, net: MotionNet, alphas=(1.0,), max_outer=2, retire=True):
        self.net = net
        self.alphas = tuple(alphas)
        self.max_outer = max_outer
        self.retire = retire
        self.reset()
        self.accepted = 0
        self.proposed = 0

    def reset(self):
        """Start of a solve: proposals resume (pyramid.solveXT calls this
        once per solve, so a proposer retired in one solve proposes again in
        the next); the accepted/proposed counters keep running."""
        self.retired = False
        self._level = None
        self._calls = 0

    def __call__(self, L, x, T, cT):
        if self._level is not L:
            self._level, self._calls = L, 0
        self._calls += 1
        active = ~cT
        none = np.zeros(len(T), bool)
        if self.retired or not active.any() or self._calls > self.max_outer:
            return T, none
        sums = lm_sums(L.y, L.E, x, T)
        # translations in voxels of the level grid
        scale = level_scale(L)
        Fh = features(sums)
        Fh[:, 0:3] /= scale
        u = propose(self.net, Fh, x.device)
        u[:, 0:3] *= scale
        d = to_update(u)
        Tn = T.copy()
        take = none.copy()
        for a in self.alphas:
            todo = active & ~take
            if not todo.any():
                break
            Tp = T + a * d
            Tp[:, 3:] = (Tp[:, 3:] + math.pi) % (2 * math.pi) - math.pi  # restrictTransform.m:28-29
            # over the readout ROI, like the LM energies in sums[:, 0]
            e_new = state_energy(L.E, x, L.y, Tp, L.off, Mbe=getattr(L, "Mbe", None))
            ok = todo & (e_new < sums[:, 0])
            Tn[ok] = Tp[ok]
            take |= ok
        self.proposed += int(active.sum())
        self.accepted += int(take.sum())
        if self.retire and not take.any():
            self.retired = True
        return Tn, take


def relative_error(T, T_true):
    """Per-shot motion error once the common motion is removed (the method
    determines the motion only up to a rigid motion shared by every shot,
    Simulation/synthesizeT.m:38, pyramidDISORDER.m:124-131): rows of
    (T - T_true) - mean over shots, translations in voxels, rotations in
    degrees."""
    d = np.asarray(T, np.float64) - np.asarray(T_true, np.float64)
    d = d - d.mean(axis=0, keepdims=True)
    d[:, 3:] *= DEG
    return d


def acceptance_curve(net, rng, dists, offsets=(0.0,), n=(64, 64, 64), S=32, C=8, n_scans=3, n_cg=4):
    """How often a proposal lowers its shot's energy (the Proposer's test),
    against the distance of the current estimate from the solution.

    For each simulated scan (simulate_scan), distance d and common offset o:
    T_c = T_true + e + c with e ~ N(0, d) per shot and parameter (voxels,
    degrees) and c ~ N(0, o) one draw shared by every shot; the image from
    n_cg CG iterations at T_c; features at (x, T_c).  Candidate updates:
      net          the proposal (propose: Newton step plus the network's
                   correction, centred over the shots),
      newton       the diagonal Newton step (features 0:6) alone.
    For each: the fraction of shots whose energy decreases, and the rms
    relative error (relative_error, all shots, 6 parameters in voxels and
    degrees) before, with every shot moved, and with only the shots that
    lowered their energy moved.  Returns a list of rows."""
    dev = net.p.device
    rows = []
    for _ in range(n_scans):
        E, x, y, T_true, off, Sm = simulate_scan(rng, n, S, C, dev)
        for o in offsets:
            for d in dists:
                e = rng.normal(0.0, d, (S, 6)) + rng.normal(0.0, o, (1, 6))
                e[:, 3:] /= DEG
                Tc = T_true + e
                xc, _, _ = ops.CGsolver(y, E, Tc, x=None, nIt=n_cg, tolType="RelativeResidual2Error", tol=0.0)
                sums = lm_sums(y, E, xc, Tc)
                F = features(sums)
                cands = {"net": to_update(propose(net, F, dev)), "newton": to_update(F[:, 0:6])}
                before = float(np.sqrt(np.mean(relative_error(Tc, T_true) ** 2)))
                row = {"dist": d, "offset": o, "err_before": before}
                for name, u in cands.items():
                    Tp = Tc + u
                    Tp[:, 3:] = (Tp[:, 3:] + math.pi) % (2 * math.pi) - math.pi
                    ok = state_energy(E, xc, y, Tp, off) < sums[:, 0]
                    Tk = Tc.copy()
                    Tk[ok] = Tp[ok]
                    row[name] = {"accept": float(ok.mean()),
                                 "err_all": float(np.sqrt(np.mean(relative_error(Tp, T_true) ** 2))),
                                 "err_accepted": float(np.sqrt(np.mean(relative_error(Tk, T_true) ** 2)))}
                rows.append(row)
                del xc
        E.close()
        del x, y, Sm
    return rows


# ------------------------------------------------------------------ training data
def simulate_scan(rng, n, S, C, dev, snr_db=30.0, sampling=False):
    """One simulated scan (the method's own synthesis): phantom, coils,
    DISORDER sampling, random-walk motion with step sigma U(0.2, 1.5) deg /
    voxel plus 5 % abrupt jumps up to 10 deg / 10 voxels, noise.  Returns
    (E, x, y, T, off, Sm), and (k1, k2) after them with `sampling`."""
    seed = int(rng.integers(1 << 30))
    t1 = int(2 ** math.floor(math.log2(math.sqrt(S))))
    t2 = S // t1
    _, k1, k2, off = synth.disorder_sampling(n[0], n[1], S, t1, t2, seed=seed)
    x = synth.phantom(n, int(rng.integers(6, 16)), seed=seed + 1, device=dev)
    Sm = synth.coil_maps(n, C, seed=seed + 2, device=dev)
    step_deg, step_vox = rng.uniform(0.2, 1.5), rng.uniform(0.2, 1.5)
    T = np.zeros((S, 6))
    for s in range(1, S):
        T[s, :3] = T[s - 1, :3] + rng.normal(0, step_vox, 3)
        T[s, 3:] = T[s - 1, 3:] + rng.normal(0, step_deg, 3) * math.pi / 180
        if rng.uniform() < 0.05:
            T[s, :3] += rng.uniform(-10, 10, 3)
            T[s, 3:] += rng.uniform(-10, 10, 3) * math.pi / 180
    T -= T.mean(axis=0, keepdims=True)
    E = ops.Encoding(Sm, k1, k2, off)
    y = E.encode(x, T)
    y = synth.add_noise(y, synth.noise_level(x, Sm, snr_db), seed=seed + 3)
    if sampling:
        return E, x, y, T, off, Sm, k1, k2
    return E, x, y, T, off, Sm


class StateRecorder:
    """A proposer for pyramid.solveXT that proposes nothing and records what
    the Proposer would see at every outer iteration of every level: the
    features at (x, T) and the relative target, translations of both in
    voxels of the level grid."""

    def __init__(self, T_true):
        self.T_true = np.asarray(T_true, np.float64)
        self.rows = []

    def reset(self):
        pass

    def __call__(self, L, x, T, cT):
        if (~cT).any():
            scale = level_scale(L)
            F = features(lm_sums(L.y, L.E, x, T))
            F[:, 0:3] /= scale
            t = relative_target(self.T_true, T)
            t[:, 0:3] /= scale
            self.rows.append((F, t))
        return T, np.zeros(len(T), bool)


def pyramid_examples(rng, n, S, C, dev, nlevels, n_outer=10, n_cg=8):
    """Training examples from the states the plain pyramid visits on one
    simulated scan (simulate_scan): solveXT with a StateRecorder, the
    default nX schedule and readout ROI, no closing CG.  Returns the rows
    and the largest relative error (voxels / degrees) of the solve's final
    motion, by which scans the plain method does not solve can be left
    out."""
    from . import pyramid
    E, x, y, T, off, Sm, k1, k2 = simulate_scan(rng, n, S, C, dev, sampling=True)
    E.close()
    rec = StateRecorder(T)
    _, Tf, _ = pyramid.solveXT(y, Sm, k1, k2, off, n, nlevels=nlevels, n_outer=n_outer, n_cg=n_cg, final_cg=0,
                               proposer=rec)
    del x, y, Sm
    return rec.rows, float(np.abs(relative_error(Tf, T)).max())


def training_examples(rng, E, x, y, T_true, off, n_cg=8, n_states=3):
    """Features and targets at n_states current estimates, with the image
    from n_cg CG iterations at that estimate (what the proposal sees inside
    solveXT).  70 % of the estimates sit near the solution (per-shot errors
    N(0, s) plus a slow walk, s ~ U(0.05, 0.6) voxel / degree: the later LM
    iterations of a level), 30 % are far (a T_true + walk, a ~ U(0, 1): the
    first iterations)."""
    S = T_true.shape[0]
    out = []
    for _ in range(n_states):
        if rng.uniform() < 0.7:
            sig = rng.uniform(0.05, 0.6)
            e = rng.normal(0, sig, (S, 6)) + np.cumsum(rng.normal(0, sig / 4, (S, 6)), axis=0)
            e[:, 3:] *= math.pi / 180
            Tc = T_true + e
        else:
            e = np.cumsum(rng.normal(0, 0.2, (S, 6)), axis=0)
            e[:, 3:] *= math.pi / 180
            Tc = rng.uniform(0.0, 1.0) * T_true + e
        xc, _, _ = ops.CGsolver(y, E, Tc, x=None, nIt=n_cg, tolType="RelativeResidual2Error", tol=0.0)
        sums = lm_sums(y, E, xc, Tc)
        out.append((features(sums), relative_target(T_true, Tc)))
    return out


def train(net: MotionNet, data, epochs=30, batch=512, lr=1e-3, seed=0, log=None, residual=True):
    """Adam on the L2 loss to the target.  data: list of (F [S x 13],
    t [S x 6]) scans; a batch holds whole scans up to `batch` rows.  With
    `residual` the network's target is net_target(F, t) (t less the Newton
    step), otherwise t itself."""
    rng = np.random.default_rng(seed)
    dev = net.p.device
    losses = []
    for ep in range(epochs):
        order = rng.permutation(len(data))
        i = 0
        tot, cnt = 0.0, 0
        while i < len(order):
            Fs, Ts, so = [], [], [0]
            while i < len(order) and (so[-1] == 0 or so[-1] + data[order[i]][0].shape[0] <= batch):
                F, t = data[order[i]]
                Fs.append(F)
                Ts.append(net_target(F, t) if residual else t)
                so.append(so[-1] + F.shape[0])
                i += 1
            X = torch.from_numpy(np.concatenate(Fs)).to(dev)
            Tg = torch.from_numpy(np.concatenate(Ts)).to(dev)
            loss = net.train_step(X, Tg, so, lr)
            tot += loss * so[-1]
            cnt += so[-1]
        losses.append(tot / max(cnt, 1))
        if log is not None:
            log(ep, losses[-1])
    return losses
