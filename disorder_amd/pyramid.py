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
  'RelativeResidual2Error', tol 0 and the nX iterations of :417-425 (:431),
  one LMsolver call on the readout ROI E.nF (:345, :507); when every state
  has converged a final 'Energy' CG closes the level (:425-431, :523);
  at the full-resolution level (AlignedRec = 2, leff = 0) the per-sample
  residuals of computeEnergy after each CG give robust per-shot weights
  (disorder_amd.robust, :441-481) that decode and the later CG calls of the
  level apply as EH.We (cleared when a level starts, :352); LMsolver and
  computeEnergy do not use them;
  the image solves encode with the motion binned by compressMotion (:392,
  disorder_amd.motion), LMsolver with the per-shot motion (:496).

  after the spatial levels, n_within full-resolution levels split the
  states of outlier shots in two (:270-287, :470-478).

Every level's data pass the Gibbs-ringing Tukey window
(downsampleOperators.m:54-57); the last `accel` levels estimate no motion
(pyramidPlan.m:23).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import motion as MB
from . import ops, robust as RB
from .sampling import level_sampling


class NotDisorder(RuntimeError):
    """A motion state without samples at some pyramid level: the sampling is
    probably not DISORDER, and solveXT.m:355 returns without performing
    corrections."""


def level_size(N, f):
    """ceil(N / 2^f) per dimension (downsampleOperators.m:34)."""
    return [int(math.ceil(v / 2 ** f)) for v in N]


def readout_roi(n3, fat_shift_dir="H", redFOV=1.0 / 3.0):
    """E.nF of Reconstruction/solveXT.m:345 as a 0-based range [lo, hi):
    fat shift 'F' keeps the first floor((1 - redFOV) n3) readout planes,
    'H' the planes floor(1 + redFOV n3) .. n3 (1-based), any other
    direction every plane (the reference then prints that performance is
    probably suboptimal)."""
    if fat_shift_dir == "F":
        return 0, int(math.floor((1.0 - redFOV) * n3))
    if fat_shift_dir == "H":
        return int(math.floor(1.0 + redFOV * n3)) - 1, int(n3)
    return 0, int(n3)


def fractional_weights(n, nF, sp, order, k1, k2):
    """E.Fd of Reconstruction/solveXT.m:353-354 at the samples: the
    'FractionalFiniteDiscreteIso' filter of Methods/buildFilter.m:58-68 on
    the grid [n1, n2, nF] (level phase-encode grid, readout ROI length) with
    spacing sp = N ./ n: k_d = 2 pi m_d / (sp_d N_d) for the FFT-order
    frequency m_d (generateGrid centred at ceil((N+1)/2), then ifftshift,
    :120), radius |k| clipped at pi, (2 |sin(|k|/2)|)^order, and the zero
    frequency set to the largest value of the rest of the grid (:134);
    buildHarmonicSampling (Methods/buildHarmonicSampling.m:35,42) takes the
    (k1, k2) row of every sample.  The largest off-zero value comes from the
    radius extreme that the order favours (the filter is monotone in the
    radius up to the clip): the grid corner for order > 0, the nearest
    neighbour of the zero frequency for order < 0.  Returns [Ns, nF]
    float32 (sample fastest when stored column-major)."""
    Nd = [int(n[0]), int(n[1]), int(nF)]
    k1 = np.asarray(k1, np.int64)
    k2 = np.asarray(k2, np.int64)
    sc = [2.0 * math.pi / (float(sp[d]) * Nd[d]) for d in range(3)]

    def freqs(m):
        i = np.arange(m)
        return np.where(i < (m + 1) // 2, i, i - m).astype(np.float64)

    def g(r):
        with np.errstate(divide="ignore"):
            return (2.0 * np.abs(np.sin(np.minimum(r, math.pi) / 2.0))) ** order

    a1 = (freqs(Nd[0]) * sc[0])[k1]
    a2 = (freqs(Nd[1]) * sc[1])[k2]
    a3 = freqs(Nd[2]) * sc[2]
    F = g(np.sqrt(a1[:, None] ** 2 + a2[:, None] ** 2 + a3[None, :] ** 2))
    live = [d for d in range(3) if Nd[d] > 1]
    if order > 0:
        dc = g(math.sqrt(sum((Nd[d] // 2 * sc[d]) ** 2 for d in live)))
    elif order < 0 and live:
        dc = g(min(sc[d] for d in live))
    else:
        dc = 1.0
    F[(k1 == 0) & (k2 == 0), 0] = dc
    return F.astype(np.float32)


def cg_iterations(li, it, sizes, nX_prev, nIt2):
    """CG iterations nX of an outer iteration while motion is estimated
    (Reconstruction/solveXT.m:417-425): nIt(2) on the first outer iteration
    of the first level and one fewer on each later one (at least 1); at a
    new resolution nIt(2) - 1, then one fewer each time; at a repeated
    resolution (the within-shot levels) none on the first outer iteration,
    nIt(2) - 2 on the second, then nX - (it - 2).  li: level (0-based),
    it: outer iteration (1-based), sizes: level grids."""
    if li == 0 and it == 1:
        return nIt2
    if li == 0:
        return max(nX_prev - 1, 1)
    same = list(sizes[li]) == list(sizes[li - 1])
    if not same and it == 1:
        return max(nIt2 - 1, 1)
    if not same:
        return max(nX_prev - 1, 1)
    if it == 1:
        return 0
    if it == 2:
        return max(nIt2 - 2, 1)
    return max(nX_prev - (it - 2), 1)


def pyramid_levels(N, nlevels):
    """[(n_l, resIso_l)] from coarse to fine (pyramidPlan.m:19-25, 1 mm voxels)."""
    out = []
    for l in range(nlevels):
        f = nlevels - 1 - l
        out.append((level_size(N, f), math.sqrt(2.0) * 2.0 ** f))
    return out


class Level:
    """Operators of one pyramid level, resident on the device.

    Coil subsets are prefixes of the (compressed, energy-ordered) coil axis,
    so each is a view of the level's maps and data.  E and y are the LMsolver
    pair (lm_coils); image solves ask cg(ncoil, T) for their encoding,
    motion, data and preconditioner."""

    def __init__(self, y, S, k1, k2, off, N, n, ctx, lm_coils=None, rows0=None, gibbs=0.0, NY=None, ny=None,
                 nF=None, fraction_order=0.0):
        self.n = list(n)
        # acquired phase-encode grid (full and at this level); below the
        # image grid the level's encodings fold (SENSE, E.Uf / EH.Ub)
        NY = list(N[:2]) if NY is None else [int(v) for v in NY[:2]]
        ny = list(n[:2]) if ny is None else [int(v) for v in ny[:2]]
        rows, l1, l2, loff = level_sampling(k1, k2, off, NY, ny)
        self.rows, self.off = rows, loff
        scale = math.sqrt(ny[0] * ny[1] / (NY[0] * NY[1]))
        C = int(S.shape[3])
        g = ops.gather_samples(y, rows if rows0 is None else np.asarray(rows0)[rows], scale, ctx)
        self.yall = ops.resampling(g, (len(rows), n[2], C), ctx) if n[2] != N[2] else g
        if gibbs:  # Gibbs-ringing filter at the level's size (downsampleOperators.m:54-57)
            ops.filter_rows(self.yall, l1, l2, (ny[0], ny[1], n[2]), gibbs, ctx)
        self.Sall = ops.resampling(S, n, ctx) if list(n) != list(N) else S
        self.ny = None if ny == list(n[:2]) else ny
        self.k1, self.k2, self.N, self.ctx = l1, l2, N, ctx
        self.C = C
        self.We = None
        self._enc = {}     # ncoil -> per-shot encoding
        self._binned = {}  # ncoil -> (offsets, encoding) of the last binning
        self._Se = {}
        lm = C if lm_coils is None else lm_coils
        self.E, self.y = self.encoding(lm), self.data(lm)
        # readout ROI of the motion estimation (E.nF) and the readout weights
        # of the energies (EH.Mbe, solveXT.m:345-346)
        self.nF = None if nF is None or tuple(nF) == (0, n[2]) else (int(nF[0]), int(nF[1]))
        self.E.set_readout_roi(self.nF)
        if fraction_order:  # E.Fd of the LMsolver pair (solveXT.m:353-354, :506)
            if self.ny is not None:
                raise ValueError("fractional finite-difference motion estimation needs unfolded data "
                                 "(buildHarmonicSampling reshapes the image-grid filter onto the acquired grid)")
            lo, hi = (0, n[2]) if self.nF is None else self.nF
            sp = [N[d] / n[d] for d in range(3)]
            self.E.set_lm_filter(fractional_weights(n, hi - lo, sp, fraction_order, l1, l2))
        self.Mbe = None
        if self.nF is not None:
            self.Mbe = np.zeros(n[2], np.float32)
            self.Mbe[self.nF[0]:self.nF[1]] = 1.0

    def _check(self, ncoil):
        if not 1 <= ncoil <= self.C:
            raise ValueError(f"coil count {ncoil} outside 1..{self.C}")

    def data(self, ncoil):
        self._check(ncoil)
        return self.yall[..., :ncoil]

    def encoding(self, ncoil):
        self._check(ncoil)
        if ncoil not in self._enc:
            self._enc[ncoil] = ops.Encoding(self.Sall[..., :ncoil], self.k1, self.k2, self.off, Nfull=self.N,
                                            We=self.We, ctx=self.ctx, NY=self.ny)
        return self._enc[ncoil]

    def precond(self, ncoil):
        """A.Se = 1 / (sum_c |S_c|^2 + 1e-9) over the first ncoil coils."""
        self._check(ncoil)
        if ncoil not in self._Se:
            self._Se[ncoil] = ops.coil_precond(self.Sall[..., :ncoil], 1e-9, self.ctx)
        return self._Se[ncoil]

    def set_weights(self, We):
        """EH.We per sample, for every encoding of the level."""
        self.We = We
        for E in list(self._enc.values()) + [e for _, e in self._binned.values()]:
            E.set_weights(We)

    def cg(self, ncoil, T, res=None, traLimX=None):
        """(encoding, motion, data, preconditioner) for an image solve over
        the first ncoil coils; with res, shots binned by bin_motion (the
        per-shot encoding when every bin holds one shot)."""
        if res is None:
            return self.encoding(ncoil), T, self.data(ncoil), self.precond(ncoil)
        Tb, offb, _ = MB.bin_motion(T, self.off, res, traLimX)
        if len(offb) == len(self.off):
            return self.encoding(ncoil), Tb, self.data(ncoil), self.precond(ncoil)
        prev = self._binned.get(ncoil)
        if prev is None or not np.array_equal(prev[0], offb):
            if prev is not None:
                prev[1].close()
            self._binned[ncoil] = (offb, ops.Encoding(self.Sall[..., :ncoil], self.k1, self.k2, offb,
                                                      Nfull=self.N, We=self.We, ctx=self.ctx, NY=self.ny))
        return self._binned[ncoil][1], Tb, self.data(ncoil), self.precond(ncoil)

    def close(self):
        for _, E in self._binned.values():
            E.close()
        for E in self._enc.values():
            E.close()
        self._binned, self._enc = {}, {}


def solveXT(y, S, k1, k2, off, N, nlevels=4, n_outer=10, n_cg=8, winit=1e-3, traLimXT=(0.05, 0.02),
            echoesMin=4, final_cg=0, T0=None, x0=None, callback=None, proposer=None, ctx=None,
            robust=True, bin_motion=True, traLimX=MB.TRA_LIM_X, eivaS=None, n_within=0, traj=None,
            gibbs=0.05, accel=0, NY=None, cg_schedule="nX", fat_shift_dir="H", redFOV=1.0 / 3.0,
            tolerSolve=1e-3, fraction_order=0.0, closing_probe=None):
    """Returns (x [N] on the device, T [states x 6], info).  callback(level,
    it, x, T, t_solve) runs after every outer iteration (and after a level's
    final CG, with outer = last + 1) outside the timed region
    (t_solve: solve seconds so far; callback time is reported in
    info['callback_s'] and excluded from info['solve_s']).
    proposer(level_ops, x, T, cT) -> (T', take) may propose motion before
    each LM call: states in `take` kept a proposal that lowered their energy
    (computeEnergy) and skip this LM call; LMsolver runs on the others, so
    the plain method stays the fallback.  robust: robust per-shot weights
    at the full-resolution levels (info['We'] holds the last ones).
    bin_motion: the image solves and the energies for the weights use the
    binned motion (disorder_amd.motion.bin_motion at the resolution of the
    finest level); LMsolver keeps the per-shot motion.
    eivaS: coil counts (final image, motion estimation, image solves
    while estimating motion) as compressCoils returns them for three
    energy fractions (solveXT.m:401,501): CG calls use the first eivaS[2]
    coils, the closing CG of a level the first eivaS[0], LMsolver (and a
    proposer) the first eivaS[1]; None uses every coil everywhere.
    n_within: full-resolution levels after the spatial pyramid
    (round(log2 NWend), pyramidPlan.m:21-22).  At each of them a level that
    converges closes with the final CG, then robust weights over the
    15..85 % percentiles flag the shots with weight below one and the next
    level splits each of their states at its middle profile in acquisition
    order (solveXT.m:270-287, :470-478); after a split only the new states
    are estimated on the first outer iteration (every iteration before the
    last level, :383-385).  Needs traj, the dict of
    disorder_amd.trajectory.trajectory_states for the scan (its state map
    must give k1, k2, off); the returned T then has one row per final state
    and info['sweep_of_state'] maps states to shots.
    gibbs: parXT.UseGiRingi * parXT.GibbsRingi (default 0.05): every
    level's data pass through the radial Tukey window of that width
    (downsampleOperators.m:54-57); 0 switches the filter off.
    accel: the last `accel` levels estimate no motion (pyramidPlan.m:23,
    parXT.accel(1)): such a level runs only its closing CG (final_cg
    iterations, 'Energy' 1e-3), with the last robust weights when leff > 0
    (solveXT.m:412-414), and flags shots for a further level as converged
    levels do.
    NY: the acquired phase-encode grid (size(y) in solveXT.m:69) when it is
    smaller than N (SENSE folding, E.Uf / EH.Ub of solveXT.m:198-200,
    :362-364): k1/k2 index it, each level's data live on ceil(NY/2^l)
    (downsampleOperators.m:34) and its encodings fold the level image onto
    that grid.
    cg_schedule: 'nX' (default) runs the CG iterations of solveXT.m:417-425
    with nIt(2) = n_cg (cg_iterations; no CG on the first outer iteration
    of a repeated full-resolution level), 'fixed' runs n_cg every time.
    fat_shift_dir ('H' by default, IO/disorderAlgorithm.m:71) and redFOV
    (1/3, :97): the readout ROI E.nF of every LMsolver call and the readout
    weights EH.Mbe of the energies behind the robust weights
    (solveXT.m:345-346, :443); a direction other than 'F' or 'H' keeps
    every readout plane.
    tolerSolve: the 'Energy' tolerance of the closing CG of a level
    (parXT.tolerSolve, IO/disorderAlgorithm.m:108; solveXT.m:177).
    fraction_order: parXT.fractionOrder (0, off, by default,
    IO/disorderAlgorithm.m:114): nonzero gives every LMsolver call the
    fractional finite-difference weights E.Fd of that order
    (fractional_weights, solveXT.m:353-354, :506); needs unfolded data.
    closing_probe(level, run, t_solve): called outside the timed region
    right before every closing CG; run(m) returns, as a new full-size
    image, the result of the first m iterations of that closing CG (the
    same operator, weights and preconditioner, from the same image; CG's
    iterates do not depend on its stop rule; run(m, energy_rule=True) does
    the closing CG's per-iteration energy checks too, for timing), so a
    caller can locate a crossing inside the closing CG without changing the
    solve.
    info['closing_cg'] lists per closing CG its level, iterations and the
    solve clock before and after it."""
    import torch
    from . import trajectory as TJ
    from .sampling import harmonic_sampling_repeats
    ctx = ctx or ops.default_context(y.device.index or 0)
    nsweeps = len(off) - 1
    T = np.zeros((nsweeps, 6)) if T0 is None else np.array(T0, np.float64, copy=True)
    x = x0
    info = {"levels": [], "callback_s": 0.0, "lm_trials": 0, "proposals_accepted": 0}
    t0 = time.perf_counter()
    cb_time = 0.0
    res_fine = pyramid_levels(N, nlevels)[-1][1]
    C = int(S.shape[3])
    ev = [C, C, C] if eivaS is None else [int(v) for v in eivaS]
    if len(ev) != 3:
        raise ValueError("eivaS holds three coil counts")
    if n_within and traj is None:
        raise ValueError("within-shot levels need the trajectory (traj)")
    if cg_schedule not in ("nX", "fixed"):
        raise ValueError("cg_schedule is 'nX' or 'fixed'")
    if proposer is not None and hasattr(proposer, "reset"):
        proposer.reset()
    # state bookkeeping for within-shot subdivision (solveXT.m:270-287)
    sweep_of_state = np.arange(nsweeps)
    k1c, k2c, offc, rows0 = k1, k2, np.asarray(off), None
    if traj is not None:
        H0 = np.asarray(traj["H"])
        H0 = H0[..., None] if H0.ndim == 2 else H0  # [NY1, NY2, repeats]
        kk1, kk2, krep, koff = harmonic_sampling_repeats(H0, traj["sweeps"])
        if not (np.array_equal(kk1, k1) and np.array_equal(kk2, k2) and np.array_equal(koff, off)):
            raise ValueError("traj does not describe the sampling k1, k2, off")
        state_of_profile = np.array(traj["sweep_of_profile"], np.int64)
        pos_row = np.full(H0.shape, -1, np.int64)
        pos_row[k1, k2, krep] = np.arange(len(k1))
        Ti = np.asarray(traj["time_index"]).reshape(H0.shape)
    subT = np.zeros(nsweeps, bool)
    subTT = np.zeros(nsweeps, bool)
    last_We = None
    if not 0 <= int(accel) <= nlevels + int(n_within):
        raise ValueError("accel must be between 0 and the number of levels")

    def cg_ops(L, T, final):
        return L.cg(ev[0] if final else ev[2], T, *((res_fine, traLimX) if bin_motion else ()))

    info["closing_cg"] = []

    def closing_cg(li, L, T, x):
        """The estT(l) = 0 CG of a level (solveXT.m:425-431, nIt(1), 'Energy' tolerSolve)."""
        nonlocal cb_time
        Ec, Tc, yc, Se = cg_ops(L, T, True)
        if closing_probe is not None:
            torch.cuda.synchronize()
            tc = time.perf_counter()

            def run(m, energy_rule=False):
                # energy_rule: the closing CG's own 'Energy' rule with tol 0 (its per-iteration work, no early stop)
                tt = ("Energy", 0.0) if energy_rule else ("RelativeResidual2Error", 0.0)
                xm = x.clone() if m <= 0 else ops.CGsolver(yc, Ec, Tc, x=x, nIt=int(m), tolType=tt[0], tol=tt[1],
                                                           Se=Se)[0]
                return xm if list(xm.shape) == list(N) else ops.resampling(xm, N, ctx)

            closing_probe(li, run, tc - t0 - cb_time)
            cb_time += time.perf_counter() - tc
            torch.cuda.synchronize()
        tb = time.perf_counter() - t0 - cb_time
        x, _, nn = ops.CGsolver(yc, Ec, Tc, x=x, nIt=final_cg, tolType="Energy", tol=tolerSolve, Se=Se)
        if closing_probe is not None:
            torch.cuda.synchronize()
        info["closing_cg"].append({"level": li, "iterations": int(nn), "t_before": tb,
                                   "t_after": time.perf_counter() - t0 - cb_time})
        return x

    levels = pyramid_levels(N, nlevels) + [(list(N), res_fine)] * int(n_within)
    last = len(levels) - 1  # L of :383, lowered to a level that flags no shot (:474)
    for li, (n, resIso) in enumerate(levels):
        leff = li - (nlevels - 1)
        if subT.any():  # split the flagged states (:270-287)
            T = np.repeat(T, np.where(subT, 2, 1), axis=0)
            subTT = np.repeat(subT, np.where(subT, 2, 1))
            sweep_of_state = np.repeat(sweep_of_state, np.where(subT, 2, 1))
            state_of_profile, _ = TJ.subdivide_states(state_of_profile, subT)
            H = np.where(Ti > 0, state_of_profile[np.maximum(Ti, 1) - 1], 0)
            k1c, k2c, repc, offc = harmonic_sampling_repeats(H, len(sweep_of_state))
            if (np.diff(offc) == 0).any():
                raise ValueError("a split left a state without samples")
            rows0 = pos_row[k1c, k2c, repc]
        subT = np.zeros(len(sweep_of_state), bool)
        nstates = len(sweep_of_state)
        ny = None if NY is None else level_size(NY[:2], max(0, nlevels - 1 - li))
        # every state keeps samples at this level, else solveXT.m:355 stops
        _, _, _, loff = level_sampling(k1c, k2c, offc, list(N[:2]) if NY is None else list(NY[:2]),
                                       list(n[:2]) if ny is None else ny)
        empty = np.flatnonzero(np.diff(loff) == 0)
        if len(empty):
            raise NotDisorder(f"level {li} ({n[0]}x{n[1]}x{n[2]}): states {empty.tolist()[:8]} have no samples; "
                              "probably not DISORDER, returning without performing corrections")
        L = Level(y, S, k1c, k2c, offc, N, n, ctx, lm_coils=ev[1], rows0=rows0, gibbs=gibbs, NY=NY, ny=ny,
                  nF=readout_roi(n[2], fat_shift_dir, redFOV), fraction_order=fraction_order)
        full_res = list(n) == list(N)
        sweep_of_sample = sweep_of_state[np.repeat(np.arange(nstates), np.diff(L.off))]
        if x is None:
            x = ops.empty_colmajor(n, y.device)
            x.zero_()
        elif list(x.shape) != list(n):
            x = ops.resampling(x, n, ctx)
        w = np.full(nstates, winit)
        cT0 = np.diff(L.off) < echoesMin
        cT = cT0.copy()
        tL = (traLimXT[0] * resIso, traLimXT[1] * resIso)
        nReset, cont = 0, 0
        its = 0
        nX = n_cg

        def weights(x, final):
            Ec, Tc, yc, _ = cg_ops(L, T, final)
            ps = ops.computeEnergy(yc, x, Ec, Tc, per_sample=True, Mbe=L.Mbe)
            We, _ = RB.shot_weights(ps, sweep_of_sample, nsweeps, finer=leff > 0)
            L.set_weights(RB.sample_weights(We, sweep_of_sample))
            info["We"] = We
            return We

        if li >= len(levels) - int(accel):  # estT(l) = 0 from the start: closing CG only
            if robust and leff > 0 and last_We is not None:
                L.set_weights(RB.sample_weights(last_We, sweep_of_sample))
            if final_cg:
                x = closing_cg(li, L, T, x)
            if robust and leff > 0:
                We = weights(x, True)
                subT = (We < 1.0)[sweep_of_state]
                if not subT.any():
                    last = li
            if callback is not None:
                torch.cuda.synchronize()
                tc = time.perf_counter()
                callback(li, 1, x, T, tc - t0 - cb_time)
                cb_time += time.perf_counter() - tc
            if li == len(levels) - 1:
                subT[:] = False
            info["levels"].append({"n": n, "outer": 0, "states_converged": nstates, "states": nstates,
                                   "split": int(subT.sum())})
            last_We = info.get("We", last_We)
            L.close()
            continue
        for it in range(1, n_outer + 1):
            if (nReset == 0 and cont == 0) or (nReset > 0 and cont % nReset == 0) or cT.all():
                nReset += 1
                cont = 0
                cT = cT0.copy()
            if subTT.any() and (it == 1 or li < last):
                cT = cT | ~subTT
            cont += 1
            Ec, Tc, yc, Se = cg_ops(L, T, False)
            nX = cg_iterations(li, it, [lv[0] for lv in levels], nX, n_cg) if cg_schedule == "nX" else n_cg
            if nX:
                x, _, _ = ops.CGsolver(yc, Ec, Tc, x=x, nIt=nX, tolType="RelativeResidual2Error", tol=0.0, Se=Se)
            if robust and full_res:
                weights(x, False)
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
                if it < n_outer:  # the estT(l) = 0 iteration of :425-481
                    if final_cg:
                        x = closing_cg(li, L, T, x)
                    if robust and leff > 0:
                        We = weights(x, True)
                        subT = (We < 1.0)[sweep_of_state]
                        if not subT.any():
                            last = li
                    if final_cg and callback is not None:
                        torch.cuda.synchronize()
                        tc = time.perf_counter()
                        callback(li, it + 1, x, T, tc - t0 - cb_time)
                        cb_time += time.perf_counter() - tc
                break
        if li == len(levels) - 1:
            subT[:] = False
        info["levels"].append({"n": n, "outer": its, "states_converged": int(cT.sum()), "states": nstates,
                               "split": int(subT.sum())})
        last_We = info.get("We", last_We)
        L.close()
    if list(x.shape) != list(N):
        x = ops.resampling(x, N, ctx)
    torch.cuda.synchronize()
    info["sweep_of_state"] = sweep_of_state
    # sampling of the final states: sample order, and rows of y for each sample
    info["sampling"] = {"k1": k1c, "k2": k2c, "off": offc,
                        "rows": np.arange(len(k1c)) if rows0 is None else np.asarray(rows0)}
    info["callback_s"] = cb_time
    info["solve_s"] = time.perf_counter() - t0 - cb_time
    return x, T, info


def _per_state(per_sample, off):
    e = np.asarray(per_sample, np.float64)
    return np.add.reduceat(e, off[:-1]) if len(e) else np.zeros(len(off) - 1)
