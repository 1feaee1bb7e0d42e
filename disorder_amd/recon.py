"""Aligned reconstruction end to end (Reconstruction/solveXT.m as called
from IO/disorderAlgorithm.m, with IO/writeData.m): coil compression, the
scan's trajectory to motion states, the joint motion + image solve over the
pyramid and the within-shot levels, the optional shearlet-regularised final
image, and the NIfTI/MAT outputs.

Input data are the scan's k-space lines on the full phase-encode grid
(y [NY1, NY2, NY3, C], column-major complex64 on the device, phase encodes
in k-space with DC first, readout in image space, zero where not acquired)
with coil maps S [N1, N2, N3, C] and the acquired phase-encode coordinates
in acquisition order (rec.Assign.z, solveXT.m:82-85).  An acquired grid
NY1 x NY2 smaller than N1 x N2 is SENSE-folded (solveXT.m:198-200).
"""
from __future__ import annotations

import numpy as np

from . import nifti, ops
from . import pyramid as PY
from . import trajectory as TJ


def aligned_reconstruction(y, S, kTraj, kRange, n_echoes, perc=(0.9, 0.9, 0.9), nlevels=4, n_outer=10, n_cg=8,
                           final_cg=300, n_within=0, accel=0, gibbs=0.05, cs=False, uncorrected_cg=0,
                           write=None, ctx=None, maximum_dynamics=4, **solve_kw):
    """Returns a dict: x (corrected image, device), T [states x 6],
    sweep_of_state, eivaS, info (of pyramid.solveXT), x_cs (with cs),
    x_uncorrected (with uncorrected_cg > 0, CG without motion, solveXT.m:242),
    files (with write = (path_out, name[, A])).

    The reference's early exits return instead {"fail": reason} and nothing
    is reconstructed (rec.Fail = 1): as many repeats as
    maximum_dynamics (parXT.maximumDynamics, disorderAlgorithm.m:117)
    or more (solveXT.m:33), a separable y-z trajectory (kTraj given as the
    pair (z2, z3) of different lengths, or one profile; :77), and a number
    of profiles that is not a whole number of shots (:83).  A level at which
    some motion state has no samples (:355, 'Probably not DISORDER') returns
    {"nowithin": True, "message": ..., "x": the reconstruction without motion
    correction (the closing CG with T = 0, :243), "T": zeros}."""
    import torch
    from .trajectory import ReconstructionFail
    N = tuple(int(v) for v in S.shape[:3])
    if y.dim() not in (4, 5) or y.shape[0] > N[0] or y.shape[1] > N[1] or y.shape[2] != N[2]:
        raise ValueError("y must be [NY1, NY2, N3, C] or [NY1, NY2, N3, C, R] with NY1 <= N1, NY2 <= N2 "
                         "(the maps' grid)")
    NY = (int(y.shape[0]), int(y.shape[1]))
    folded = NY != N[:2]
    R = int(y.shape[4]) if y.dim() == 5 else 1
    if int(y.shape[3]) != int(S.shape[3]):
        raise ValueError("y and S must hold the same coils")
    if y.dim() == 4 and not ops.is_colmajor(y):  # the rows below are read through a column-major view
        raise ValueError("y must have column-major (MATLAB) strides; use disorder_amd.colmajor(y)")
    if R >= maximum_dynamics:  # solveXT.m:33
        return {"fail": f"Problem too big to fit in memory ({R} dynamics for a limit of {maximum_dynamics})"}
    try:
        TJ.trajectory_states(kTraj, kRange, n_echoes, n_repeats=R)  # host checks of :77, :83 before device work
    except ReconstructionFail as exc:
        return {"fail": str(exc)}
    if y.dim() == 5:  # repeats next to the phase encodes, coils last: [NY1, NY2, R, N3, C]
        yp = ops.empty_colmajor((NY[0], NY[1], R, N[2], int(y.shape[3])), y.device)
        yp.copy_(y.permute(0, 1, 4, 2, 3))
        y = yp
    ctx = ctx or ops.default_context(y.device.index or 0)
    # coil compression (solveXT.m:61): maps and data on the first max(eivaS) components
    if perc is not None:
        S, y, eivaS, _, _ = ops.compressCoils(S, list(perc), y, 3, ctx)
        eivaS = [int(v) for v in np.atleast_1d(eivaS)]
        eivaS = (eivaS * 3)[:3] if len(eivaS) < 3 else eivaS[:3]
    else:
        eivaS = [int(S.shape[3])] * 3
    # trajectory -> states -> harmonic sampling (solveXT.m:75-141, :349-352)
    k1, k2, off, st = TJ.sampling_from_trajectory(kTraj, kRange, n_echoes, n_repeats=R)
    if tuple(st["H"].shape[:2]) != NY:
        raise ValueError("kRange does not span the acquired phase-encode grid of y")
    NYR = NY[0] * NY[1] * R
    C = int(y.shape[-1])
    flat = y.as_strided((NYR, N[2], C), (1, NYR, NYR * N[2]))
    rows = np.asarray(k1) + NY[0] * np.asarray(k2) + NY[0] * NY[1] * np.asarray(st["sample_repeat"], np.int64)
    yc = ops.gather_samples(flat, rows, 1.0, ctx)
    out = {"eivaS": eivaS}
    if uncorrected_cg:
        L = PY.Level(yc, S, k1, k2, off, N, N, ctx, lm_coils=eivaS[0], gibbs=gibbs, NY=NY, ny=NY)
        Ec, _, ycc, Se = L.cg(eivaS[0], np.zeros((len(off) - 1, 6)))
        out["x_uncorrected"], _, _ = ops.CGsolver(ycc, Ec, np.zeros((len(off) - 1, 6)), nIt=uncorrected_cg,
                                                  tolType="Energy", tol=1e-3, Se=Se)
        L.close()
    try:
        x, T, info = PY.solveXT(yc, S, k1, k2, off, N, nlevels=nlevels, n_outer=n_outer, n_cg=n_cg,
                                final_cg=final_cg, eivaS=eivaS, n_within=n_within, traj=st, gibbs=gibbs, accel=accel,
                                ctx=ctx, NY=NY if folded else None, **solve_kw)
    except PY.NotDisorder as exc:  # solveXT.m:355: the uncorrected reconstruction is what remains
        if "x_uncorrected" not in out:
            L = PY.Level(yc, S, k1, k2, off, N, N, ctx, lm_coils=eivaS[0], gibbs=gibbs, NY=NY, ny=NY)
            Ec, _, ycc, Se = L.cg(eivaS[0], np.zeros((len(off) - 1, 6)))
            out["x_uncorrected"], _, _ = ops.CGsolver(ycc, Ec, np.zeros((len(off) - 1, 6)), nIt=final_cg,
                                                      tolType="Energy", tol=1e-3, Se=Se)
            L.close()
        out.update({"nowithin": True, "message": str(exc), "x": out["x_uncorrected"],
                    "T": np.zeros((len(off) - 1, 6))})
        torch.cuda.synchronize()
        return out
    out.update({"x": x, "T": T, "sweep_of_state": info["sweep_of_state"], "info": info})
    if cs:
        from . import shearlet as SH
        sm = info["sampling"]
        yf = ops.gather_samples(yc, sm["rows"], 1.0, ctx)
        if gibbs:
            ops.filter_rows(yf, sm["k1"], sm["k2"], (NY[0], NY[1], N[2]), gibbs, ctx)
        sweep_of_sample = info["sweep_of_state"][np.repeat(np.arange(len(sm["off"]) - 1), np.diff(sm["off"]))]
        We = None if "We" not in info else np.asarray(info["We"])[sweep_of_sample]
        c0 = eivaS[0]
        Sf = S[..., :c0]
        E = ops.Encoding(Sf, sm["k1"], sm["k2"], sm["off"], We=We, ctx=ctx, NY=NY)
        out["x_cs"], out["lambdas"] = SH.CSsolver(yf[..., :c0], E, T, x, Sf, sm["k1"], sm["k2"], We)
        E.close()
    if write is not None:
        path_out, name = write[0], write[1]
        A = write[2] if len(write) > 2 else None
        vols = {"d": x.cpu().numpy()}
        if "x_uncorrected" in out:
            vols["x"] = out["x_uncorrected"].cpu().numpy()
        if "x_cs" in out:
            vols["r"] = out["x_cs"].cpu().numpy()
        out["files"] = nifti.write_data(path_out, name, vols, T=T, A=A)
    torch.cuda.synchronize()
    return out
