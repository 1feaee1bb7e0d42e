"""Joint motion + image reconstruction over a spatial pyramid on the GPU
(the loop of Reconstruction/solveXT.m:266-531 for isotropic voxels and one
repeat), driving the C-ABI operators level by level:

  per level l (coarse to fine, sizes ceil(N/2^f), f = L-1..0,
  Methods/downsampleOperators.m:34):
    * S_l = resampling(S, n_l) (downsampleOperators.m:48-50),
    * y_l = the compact rows whose phase-encode frequencies survive the
      k-space crop of the sampling (:36), scaled by sqrt(n1 n2 / (N1 N2)) so
      the level's unitary DFT convention holds (solveXT.m:342), with the
      readout resampled in image space (:51),
    * Se = 1/(sum_c |S_c|^2 + 1e-9) (solveXT.m:231-232),
    * x resampled from the previous level (solveXT.m:531 then :333),
    * w = winit, cT reset; states with fewer than echoesMin samples frozen
      (:313-326); tL = traLimXT * resIso(l) (:370);
  each outer iteration: motion-convergence reset schedule (:376-381), CG with
  'RelativeResidual2Error', tol 0 and n_cg iterations (:431), one LMsolver
  call (:507); when every state has converged a final 'Energy' CG closes the
  level (:425-431, :523);
  at the full-resolution level (AlignedRec = 2, leff = 0) the per-sample
  residuals of computeEnergy after each CG give robust per-shot weights
  (disorder_amd.robust, :441-481) that decode and the later CG calls of the
  level apply as EH.We (cleared when a level starts, :352); LMsolver and
  computeEnergy do not use them.

Omitted from solveXT (documented deviations): the Gibbs-ringing apodisation
of y (downsampleOperators.m:54-57), within-shot subdivision (:482-499),
compressMotion binning (:392).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import ops, robust as RB
from .sampling import level_sampling


def pyramid_levels(N, nlevels):
    """[(n_l, resIso_l)] from coarse to fine (pyramidPlan.m:19-25, 1 mm voxels)."""
    out = []
    for l in range(nlevels):
        f = nlevels - 1 - l
        out.append(([int(math.ceil(v / 2 ** f)) for v in N], math.sqrt(2.0) * 2.0 ** f))
    return out


class Level:
    """Operators of one pyramid level, resident on the device."""

    def __init__(self, y, S, k1, k2, off, N, n, ctx):
        self.n = list(n)
        rows, l1, l2, loff = level_sampling(k1, k2, off, N[:2], n[:2])
        self.rows, self.off = rows, loff
        scale = math.sqrt(n[0] * n[1] / (N[0] * N[1]))
        C = int(S.shape[3])
        g = ops.gather_samples(y, rows, scale, ctx)
        self.y = ops.resampling(g, (len(rows), n[2], C), ctx) if n[2] != N[2] else g
        self.S = ops.resampling(S, n, ctx) if list(n) != list(N) else S
        self.Se = ops.coil_precond(self.S, 1e-9, ctx)
        self.E = ops.Encoding(self.S, l1, l2, loff, Nfull=N, ctx=ctx)


def solveXT(y, S, k1, k2, off, N, nlevels=4, n_outer=10, n_cg=8, winit=1e-3, traLimXT=(0.05, 0.02),
            echoesMin=4, final_cg=0, T0=None, x0=None, callback=None, proposer=None, ctx=None,
            robust=True):
    """Returns (x [N] on the device, T [S x 6], info).  callback(level, it, x,
    T, t_solve) runs after every outer iteration (and after a level's final
    CG, with outer = last + 1) outside the timed region
    (t_solve: solve seconds so far; callback time is reported in
    info['callback_s'] and excluded from info['solve_s']).
    proposer(level_ops, x, T, cT) -> (T', take) may propose motion before
    each LM call: states in `take` kept a proposal that lowered their energy
    (computeEnergy) and skip this LM call; LMsolver runs on the others, so
    the plain method stays the fallback.  robust: robust per-shot weights
    at the full-resolution level (info['We'] holds the last ones)."""
    import torch
    ctx = ctx or ops.default_context(y.device.index or 0)
    S_states = len(off) - 1
    T = np.zeros((S_states, 6)) if T0 is None else np.array(T0, np.float64, copy=True)
    x = x0
    info = {"levels": [], "callback_s": 0.0, "lm_trials": 0, "proposals_accepted": 0}
    t0 = time.perf_counter()
    cb_time = 0.0
    for li, (n, resIso) in enumerate(pyramid_levels(N, nlevels)):
        L = Level(y, S, k1, k2, off, N, n, ctx)
        full_res = list(n) == list(N)
        st_of_sample = np.repeat(np.arange(S_states), np.diff(L.off))
        if x is None:
            x = ops.empty_colmajor(n, y.device)
            x.zero_()
        elif list(x.shape) != list(n):
            x = ops.resampling(x, n, ctx)
        w = np.full(S_states, winit)
        cT0 = np.diff(L.off) < echoesMin
        cT = cT0.copy()
        tL = (traLimXT[0] * resIso, traLimXT[1] * resIso)
        nReset, cont = 0, 0
        its = 0
        for it in range(1, n_outer + 1):
            if (nReset == 0 and cont == 0) or (nReset > 0 and cont % nReset == 0) or cT.all():
                nReset += 1
                cont = 0
                cT = cT0.copy()
            cont += 1
            x, _, _ = ops.CGsolver(L.y, L.E, T, x=x, nIt=n_cg, tolType="RelativeResidual2Error", tol=0.0, Se=L.Se)
            if robust and full_res:
                ps = ops.computeEnergy(L.y, x, L.E, T, per_sample=True)
                We, _ = RB.shot_weights(ps, st_of_sample, S_states, finer=False)
                L.E.set_weights(RB.sample_weights(We, st_of_sample))
                info["We"] = We
            if proposer is not None:
                T_before = T.copy()
                T, take = proposer(L, x, T, cT)
                info["proposals_accepted"] += int(take.sum())
                T, w, cT_lm, _, ntr = ops.LMsolver(L.y, L.E, x, T, w, cT | take, winit=winit, tL=tL)
                d = np.abs(T - T_before)
                small = (d[:, :3].max(axis=1) < tL[0]) & (d[:, 3:].max(axis=1) * 180 / math.pi < tL[1])
                cT = np.where(take, small, cT_lm)
            else:
                T, w, cT, _, ntr = ops.LMsolver(L.y, L.E, x, T, w, cT, winit=winit, tL=tL)
            info["lm_trials"] += ntr
            its = it
            if callback is not None:
                torch.cuda.synchronize()
                tc = time.perf_counter()
                callback(li, it, x, T, tc - t0 - cb_time)
                cb_time += time.perf_counter() - tc
            if cT.all():
                if final_cg and it < n_outer:
                    x, _, _ = ops.CGsolver(L.y, L.E, T, x=x, nIt=final_cg, tolType="Energy", tol=1e-3, Se=L.Se)
                    if callback is not None:
                        torch.cuda.synchronize()
                        tc = time.perf_counter()
                        callback(li, it + 1, x, T, tc - t0 - cb_time)
                        cb_time += time.perf_counter() - tc
                break
        info["levels"].append({"n": n, "outer": its, "states_converged": int(cT.sum())})
        L.E.close()
    if list(x.shape) != list(N):
        x = ops.resampling(x, N, ctx)
    torch.cuda.synchronize()
    info["callback_s"] = cb_time
    info["solve_s"] = time.perf_counter() - t0 - cb_time
    return x, T, info


def _per_state(per_sample, off):
    e = np.asarray(per_sample, np.float64)
    return np.add.reduceat(e, off[:-1]) if len(e) else np.zeros(len(off) - 1)
