"""Robust per-shot weights of the aligned reconstruction
(Reconstruction/solveXT.m:450-481, defaults of solveXT.m parXT:
percRobustShot = [0.125 0.25], enerRobustShot = 0.95).

From the per-sample residual energies of computeEnergy (one value per
sampled (k1, k2) line, summed over readout and coils) each shot gets the
percentiles of the logs of its samples' residuals (15:5:85 % at the
motion-free levels, 30:5:70 % otherwise, :452-455).  Across shots, per
percentile, a robust standard deviation (the spread between the 12.5 % and
25 % percentiles over the matching normal quantiles, :458) and a robust mean
(the 18.75 % percentile moved by that many deviations back to the centre,
:459-460) give a z-score; each shot's mean z-score z maps to the weight
min((1 - Phi(z)) / ((1 - 0.95) / NSweeps), 1) (:461-464): shots whose
residuals sit well above the others get weights below one, and decode
(E^H We y) and the CG normal equations use them (decode.m:38-46).

Percentiles follow MATLAB's prctile: the i-th smallest of n values sits at
100 (i - 0.5) / n, with linear interpolation between and clamping outside
(numpy's 'hazen' method).  The work is a few thousand values per call, so it
runs on the host between the device's energy pass and the next CG.
"""
from __future__ import annotations

import numpy as np
from scipy import special

PERC_ROBUST_SHOT = (0.125, 0.25)
ENER_ROBUST_SHOT = 0.95


def percent_use(finer: bool) -> np.ndarray:
    """solveXT.m:452-454: 15:5:85 when leff > 0, else 30:5:70 (percent)."""
    return (np.arange(15, 90, 5) if finer else np.arange(30, 75, 5)).astype(np.float64)


def prctile(x, p, axis=None):
    """MATLAB prctile (Hazen plotting positions, clamped at the ends)."""
    return np.percentile(np.asarray(x, np.float64), p, axis=axis, method="hazen")


def shot_weights(per_sample, state_of_sample, nshots: int, finer: bool = False,
                 perc=PERC_ROBUST_SHOT, ener: float = ENER_ROBUST_SHOT):
    """Weights We[nshots] in (0, 1] from per-sample residual energies.

    per_sample: [nsamples] residual energy of each sampled line;
    state_of_sample: [nsamples] shot index of each line.  A shot without
    lines has no percentiles and stays out of the statistics across shots,
    as prctile skips NaN rows, and gets weight 1 (it weights nothing);
    NSweeps in the weight still counts it.  Returns (We, z) with z the mean
    robust z-score (NaN for such shots)."""
    r = np.asarray(per_sample, np.float64)
    st = np.asarray(state_of_sample, np.int64)
    if r.shape != st.shape:
        raise ValueError("per_sample and state_of_sample differ in length")
    if nshots < 1:
        raise ValueError("nshots must be positive")
    if len(st) and (st.min() < 0 or st.max() >= nshots):
        raise ValueError("state_of_sample out of range")
    counts = np.bincount(st, minlength=nshots)
    live = np.flatnonzero(counts)
    if len(live) == 0:
        raise ValueError("no samples")
    pu = percent_use(finer)
    logr = np.log(r)
    order = np.argsort(st, kind="stable")
    bounds = np.concatenate([[0], np.cumsum(counts)])
    grouped = logr[order]
    WeO = np.stack([prctile(grouped[bounds[s]:bounds[s + 1]], pu) for s in live])  # [live shots, perc]
    p = np.asarray(perc, np.float64)
    q = prctile(WeO, p * 100.0, axis=0)  # [2, perc]
    std = (q[1] - q[0]) / (special.ndtri(p[1]) - special.ndtri(p[0]))
    pm = float(p.mean())
    mea = prctile(WeO, pm * 100.0, axis=0) + std * np.sqrt(2.0) * special.erfcinv(2.0 * pm)
    z = np.full(nshots, np.nan)
    z[live] = ((WeO - mea) / std).mean(axis=1)
    We = np.ones(nshots)
    We[live] = np.minimum((1.0 - special.ndtr(z[live])) / ((1.0 - ener) / nshots), 1.0)
    return We, z


def sample_weights(We, state_of_sample):
    """EH.We per sample from the shot weights (solveXT.m:479)."""
    return np.asarray(We, np.float64)[np.asarray(state_of_sample, np.int64)]
