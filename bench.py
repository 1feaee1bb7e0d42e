"""Benchmark: E^H E applications per second on one MI355X.

Workload: the encoding operator of DISORDER's reconstruction path at
256x192x256 (phase encode 256x192, readout 256), 8 coils, 64 DISORDER shots
(8x8 random-checkered tiles, 768 phase-encode lines per shot), random-walk
rigid motion per shot, FP32 complex.  One step = one E^H We E x
(dis_encodedecode) with every input already resident in HBM.

Prints one JSON line.  ``roofline`` is measured live with HIP events recorded
by the library around every launch of the dominant kernel during the timed
steps; ``cpu_baseline`` times the FP64 CPU oracle on a bounded sample (one
shot's E^H E contribution) on this host.  The line also carries the median
of 20 single applications and the joint pyramid's time to 1e-3 NRMSE.

Other workloads (--workload): pyramid (the joint CG+LM reconstruction, plain
and with the learned proposal), compress (64 -> 8 coil compression),
batched (six uneven volumes, hipGraph-captured CG), cs (the shearlet
regulariser), small (the 64^3 reference case end to end, with the oracle's
run beside it), fold (E^H E with SENSE folding of the phase encodes), stream
(the device-copy ceiling), proposal (how often the learned proposal
lowers a shot's energy, against the distance from the solution) and lm (one
LMsolver call and its partial derivatives at 256x192x256, plain and with
fractional finite-difference weights).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pyramid", action="store_true",
                    help="E^H E line without the time-to-NRMSE summary of the joint pyramid (plain and learned)")
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--cpu-full", dest="cpu_full", action="store_true",
                    help="--workload small: run the oracle's whole 64^3 case (minutes), not a bounded sample")
    ap.add_argument("--workload", choices=["ehe", "pyramid", "compress", "batched", "cs", "small", "fold", "stream",
                                           "proposal", "lm"],
                    default="ehe",
                    help="ehe: E^H E apps/s (default); pyramid: time to 1e-3 NRMSE of the joint CG+LM pyramid; "
                         "compress: 64 -> 8 coil compression of maps and data; batched: six uneven volumes in one "
                         "batched CG with hipGraph-captured iterations; small: the 64^3 reference case (CG with "
                         "known motion, then the 3-level pyramid); fold: E^H E with SENSE folding; stream: device-copy "
                         "ceiling; proposal: learned-proposal acceptance against distance from the solution; lm: "
                         "LMsolver and its partials at 256x192x256")
    ap.add_argument("--plain-only", dest="plain_only", action="store_true",
                    help="pyramid workload: one plain solve only, without the reference and learned runs")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--outer", type=int, default=10)
    ap.add_argument("--cg", type=int, default=8)
    ap.add_argument("--final-cg", dest="final_cg", type=int, default=300,
                    help="CG iterations (Energy, tol 1e-3) closing a level whose states all converged (nIt(1))")
    ap.add_argument("--learned", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "disorder_amd",
                                                      "weights", "motion_net.npy"),
                    help="weights of the learned proposal for --workload pyramid ('none' to skip)")
    return ap.parse_args()


def cpu_baseline(n, C, S, seed):
    """The FP64 oracle (scipy.fft on all host cores) on one shot's E^H E
    contribution at the full grid; apps/s = 1 / (S * t_shot)."""
    import numpy as np
    from oracle import disorder_oracle as O
    from disorder_amd import sampling

    workers = min(os.cpu_count() or 1, 16)
    O.set_fft_workers(workers)
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex128)
    Sm = (rng.standard_normal(n + (C,)) + 1j * rng.standard_normal(n + (C,))) * 0.3
    perms = sampling.permutation_table(rng, (n[0] // 8) * (n[1] // 8), S)
    H = sampling.random_checkered(n[0], n[1], 8, 8, perms)
    H1 = np.where(H == 1, 1, 0)  # the sample set of one shot
    k1, k2, off = sampling.harmonic_sampling(H1, 1)
    samp = O.Sampling(k1, k2, off, n[0], n[1])
    T = np.array([[1.2, -0.7, 0.4, 0.03, -0.05, 0.02]])
    t0 = time.perf_counter()
    O.encodedecode(x, Sm, samp, T, n)
    dt = time.perf_counter() - t0
    O.set_fft_workers(None)
    # the reference's own algorithm: every DFT a GEMM with single(dftmtx(n))
    # (fftGPU.m), on the host BLAS threads
    O.set_dense_dft(True)
    t0 = time.perf_counter()
    O.encodedecode(x, Sm, samp, T, n)
    dd = time.perf_counter() - t0
    O.set_dense_dft(False)
    blas = int(os.environ.get("OPENBLAS_NUM_THREADS") or os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {"value": 1.0 / (S * dt), "unit": "E^H E apps/s", "cores": workers, "kind": "port",
            "sample": f"FP64 oracle (scipy.fft, {workers} threads) E^H E of 1 of {S} shots at "
                      f"{n[0]}x{n[1]}x{n[2]}, {C} coils: {dt:.2f} s; apps/s = 1/({S} x that)",
            "dense_dft_fp32": {"value": 1.0 / (S * dd), "unit": "E^H E apps/s", "cores": blas,
                               "sample": f"the same shot with every DFT a complex64 GEMM by dftmtx(n) as "
                                         f"Methods/fftGPU.m applies it (numpy tensordot on the host BLAS, "
                                         f"{blas} threads): {dd:.2f} s"}}


def reconstruction_problem(dev, n=(256, 192, 256), C_in=32, C=8, S=64, snr_db=25.0, seed=0):
    """Brain-scale plain-method problem: 40 nested ellipsoids with smooth phase,
    C_in Gaussian coils compressed to C by compressCoils, DISORDER 8x8
    random-checkered sampling, random-walk motion (0.5 deg / 0.5 voxel steps,
    clipped at 8), complex Gaussian noise at snr_db (SNRToLevels.m)."""
    import disorder_amd as D
    from disorder_amd import synth
    t = int(round(math.sqrt(S)))
    _, k1, k2, off = synth.disorder_sampling(n[0], n[1], S, t, S // t, seed=seed + 1)
    x_gt = synth.phantom(n, 40, seed=seed + 2, device=dev)
    S_in = synth.coil_maps(n, C_in, seed=seed + 3, device=dev)
    T_gt = synth.random_walk_motion(S, seed=seed + 4)
    E_in = D.Encoding(S_in, k1, k2, off)
    y = E_in.encode(x_gt, T_gt)
    E_in.close()
    y = synth.add_noise(y, synth.noise_level(x_gt, S_in, snr_db), seed=seed + 5)
    Sc, yc, cc, _, _ = D.compressCoils(S_in, C, y)
    del S_in, y
    return {"x_gt": x_gt, "S": Sc, "y": yc, "k1": k1, "k2": k2, "off": off, "T_gt": T_gt, "n": n, "C": C, "S_states": S}


def _closing_crossing(run, x_ref, nref, m_start, m_cap, tol=1e-3):
    """Smallest even m (to 2 iterations) in (0, m_cap] with the closing CG's
    m-th image within tol NRMSE of x_ref, by doubling from m_start and
    bisecting; None when even m_cap is not.  run(m) replays m iterations.
    Returns (m or None, {m: NRMSE})."""
    import torch
    seen = {}

    def err(m):
        if m not in seen:
            seen[m] = torch.linalg.vector_norm(run(m) - x_ref).item() / nref
        return seen[m]

    lo, hi = 0, max(2, min(int(m_start), m_cap))
    while err(hi) > tol:
        if hi >= m_cap:
            return None, seen
        lo, hi = hi, min(2 * hi, m_cap)
    while hi - lo > 2:
        mid = (lo + hi) // 2
        if err(mid) <= tol:
            hi = mid
        else:
            lo = mid
    return hi, seen


def _run_pyramid(P, args, proposer=None, x_ref=None, tolerSolve=1e-3, label=None, probe_start=16, probe_cap=None):
    """One joint reconstruction with per-outer-iteration snapshots (taken
    outside the timed solve): NRMSE against the ground truth after removing
    the common motion (the image moved by mean(T - T_gt), as
    Simulation/pyramidDISORDER.m:124-131 removes it) and against x_ref.
    With x_ref, the finest level's closing CG is replayed from its starting
    image (outside the timed solve) to find the first iteration m* whose
    image lies within 1e-3 NRMSE of x_ref; the time to 1e-3 is then the
    solve clock before that closing CG plus the time of that closing CG
    run to iteration m* with its own per-iteration energy checks (timed on
    a replay; m* may exceed the iterations the run's own stop rule made:
    the time the plain method takes when its closing CG goes on).
    Returns (x, T, info, curve, crossing)."""
    import numpy as np
    import torch
    import disorder_amd as D
    from disorder_amd import pyramid as PY
    n = P["n"]
    xg = P["x_gt"]
    ng = torch.linalg.vector_norm(xg).item()
    nref = torch.linalg.vector_norm(x_ref).item() if x_ref is not None else None
    snaps = []
    name = label or ("learned" if proposer else "plain")

    def aligned(xf, T):
        d = np.mean(np.asarray(T) - P["T_gt"], axis=0, keepdims=True)
        xa = D.sincRigidTransform(xf, d, di=1)[..., 0]
        return torch.linalg.vector_norm(xa - xg).item() / ng

    def cb(level, it, x, T, t_solve):
        xf = D.resampling(x, n) if list(x.shape) != list(n) else x.clone()
        e_ref = torch.linalg.vector_norm(xf - x_ref).item() / nref if x_ref is not None else None
        snaps.append({"level": level, "outer": it, "t_s": t_solve, "nrmse_vs_gt_aligned": aligned(xf, T),
                      "nrmse_vs_ref": e_ref})
        print(f"[pyramid] {name} level {level} outer {it}: solve {t_solve:.2f} s, NRMSE vs GT (aligned) "
              f"{snaps[-1]['nrmse_vs_gt_aligned']:.5f}" + (f", vs reference {e_ref:.2e}" if e_ref is not None else ""),
              file=sys.stderr, flush=True)

    crossing = {}

    def probe(level, run, t_solve):
        if x_ref is None or level != args.levels - 1:
            return
        m, seen = _closing_crossing(run, x_ref, nref, probe_start, probe_cap or args.final_cg)
        crossing.update({"m": m, "t_before_s": t_solve, "nrmse_vs_ref_at": {str(k): v for k, v in sorted(seen.items())}})
        if m is not None:  # the closing CG's time to its m-th iteration, with its own per-iteration energy checks
            run(m, energy_rule=True)
            torch.cuda.synchronize()
            tr = time.perf_counter()
            run(m, energy_rule=True)
            torch.cuda.synchronize()
            crossing["closing_s_to_m"] = time.perf_counter() - tr

    x, T, info = PY.solveXT(P["y"], P["S"], P["k1"], P["k2"], P["off"], n, nlevels=args.levels,
                            n_outer=args.outer, n_cg=args.cg, final_cg=args.final_cg, callback=cb, proposer=proposer,
                            tolerSolve=tolerSolve, closing_probe=probe)
    if crossing:
        cc = [c for c in info["closing_cg"] if c["level"] == args.levels - 1][-1]
        crossing["closing_iterations"] = cc["iterations"]
        crossing["closing_s"] = cc["t_after"] - cc["t_before"]
        early = next((sn["t_s"] for sn in snaps if sn["nrmse_vs_ref"] is not None and sn["nrmse_vs_ref"] <= 1e-3
                      and sn["t_s"] <= cc["t_before"] + 1e-9), None)
        crossing["time_s"] = early if early is not None else (
            cc["t_before"] + crossing["closing_s_to_m"] if crossing["m"] is not None else None)
    return x, T, info, snaps, crossing


def bench_pyramid(args):
    """Joint reconstruction at 256x192x256 (32 coils compressed to 8, 64
    shots, 4 levels, 10 LM iterations per level, the nX CG schedule of
    solveXT.m:417-425 with nIt(2) = 8, readout ROI 'H', a level closed by CG
    to 'Energy' tol 1e-3 once every state converged), plain and with the
    learned proposal.  value = the plain method's solve time to converged
    motion and its closing CG.  Reference: the plain method run with the
    closing CG at 'Energy' tol 1e-6 (untimed); each run reports the time at
    which its image first lies within 1e-3 NRMSE of it: the solve clock
    before the finest level's closing CG plus the time that closing CG
    needs to get there (its iteration found by replaying it from its
    starting image, outside the timed solve, and the time of a replay to
    that iteration with the closing CG's own energy checks).  The default
    closing CG stops at 'Energy' 1e-3, before that point
    (final_nrmse_vs_ref says where), so the plain time counts the
    iterations its closing CG would go on for; the learned run is searched
    up to twice the plain method's iterations and is null when its image
    converges elsewhere.  The learned run's solve-time ratio is also taken
    at its own convergence, with the NRMSE between the two final images
    (the criterion of equal quality: the plain NRMSE within 1e-4).
    NRMSE against the ground truth is given frame-aligned, with the motion
    error once the common motion is removed.  Snapshots are taken outside
    the timed solve."""
    import numpy as np
    import torch
    from disorder_amd import learned as LN

    dev = torch.device("cuda", 0)
    torch.cuda.synchronize()
    tg = time.perf_counter()
    P = reconstruction_problem(dev, S=args.shots)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - tg
    n = P["n"]

    def summary(x, T, info, snaps):
        T_err = P["T_gt"] - T
        T_err = T_err - T_err.mean(axis=0, keepdims=True)
        return {"solve_s": info["solve_s"], "levels": info["levels"], "lm_trials": info["lm_trials"],
                "proposals_accepted": info["proposals_accepted"],
                "final_nrmse_vs_gt_aligned": snaps[-1]["nrmse_vs_gt_aligned"],
                "final_nrmse_vs_ref": snaps[-1]["nrmse_vs_ref"],
                "motion_err_max_vox": float(np.abs(T_err[:, :3]).max()),
                "motion_err_max_deg": float(np.abs(T_err[:, 3:]).max() * 180 / math.pi),
                "motion_err_rms_vox": float(np.sqrt(np.mean(T_err[:, :3] ** 2))),
                "motion_err_rms_deg": float(np.sqrt(np.mean(T_err[:, 3:] ** 2)) * 180 / math.pi)}

    if args.plain_only:
        # one timed plain solve and nothing else (for kernel traces of the solve)
        xp, Tp, ip, sp, _ = _run_pyramid(P, args)
        return {"metric": "joint CG+LM pyramid solve time to converged motion (plain LM), s", "value": ip["solve_s"],
                "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "complex64",
                "data": "synthetic (seeded phantom, 32 Gaussian coils compressed to 8, DISORDER 8x8 sampling, "
                        "random-walk motion, 25 dB noise)",
                "config": {"workload": f"joint CG+LM pyramid at {n[0]}x{n[1]}x{n[2]}, 8 coils, {args.shots} shots, "
                                       f"{args.levels} levels, plain solve only"},
                "data_gen_s": gen_s, "plain": summary(xp, Tp, ip, sp)}
    xr, Tr, ir, sr, _ = _run_pyramid(P, args, tolerSolve=1e-6, label="reference")
    reference = summary(xr, Tr, ir, sr)
    xp, Tp, ip, sp, cp = _run_pyramid(P, args, x_ref=xr)
    plain = summary(xp, Tp, ip, sp)
    plain["time_to_1e-3_vs_ref_s"] = cp.get("time_s")
    plain["crossing"] = cp
    plain["curve"] = sp
    # known-motion CG on the same data for comparison (closing CG settings)
    import disorder_amd as D
    E = D.Encoding(P["S"], P["k1"], P["k2"], P["off"])
    Se = D.coil_precond(P["S"])
    xk, _, nk = D.CGsolver(P["y"], E, P["T_gt"], nIt=args.final_cg, tolType="Energy", tol=1e-3, Se=Se)
    E.close()
    known = torch.linalg.vector_norm(xk - P["x_gt"]).item() / torch.linalg.vector_norm(P["x_gt"]).item()
    del xk, Se
    line = {"metric": "joint CG+LM pyramid solve time to converged motion (plain LM), s", "value": ip["solve_s"],
            "unit": "s", "n_gpus": 1,
            "higher_is_better": False, "dtype": "complex64",
            "data": "synthetic (seeded phantom, 32 Gaussian coils compressed to 8, DISORDER 8x8 sampling, "
                    "random-walk motion, 25 dB noise)",
            "config": {"workload": f"joint CG+LM pyramid at {n[0]}x{n[1]}x{n[2]}, 8 coils, {args.shots} shots, "
                                   f"{args.levels} levels, {args.outer} outer, nX schedule nIt(2) = {args.cg}, "
                                   f"readout ROI 'H', final CG {args.final_cg} (Energy 1e-3)",
                       "reference": "the plain method with its closing CG at Energy 1e-6 (untimed)"},
            "data_gen_s": gen_s, "plain": plain, "reference": reference,
            "known_motion_cg_nrmse_vs_gt": known, "known_motion_cg_iterations": nk}
    wpath = args.learned if args.learned not in (None, "", "none") else None
    if wpath is not None and os.path.exists(wpath):
        prop = LN.Proposer(LN.MotionNet.load(wpath))
        # the learned run's closing CG is searched up to twice the plain method's crossing
        mp = cp.get("m") or args.final_cg
        xl, Tl, il, sl, cl = _run_pyramid(P, args, proposer=prop, x_ref=xr, probe_start=mp,
                                          probe_cap=min(args.final_cg, 2 * mp))
        learned = summary(xl, Tl, il, sl)
        learned["time_to_1e-3_vs_ref_s"] = cl.get("time_s")
        learned["crossing"] = cl
        learned["proposals"] = {"proposed": prop.proposed, "accepted": prop.accepted, "retired": prop.retired}
        learned["curve"] = sl
        line["learned"] = learned
        line["learned_nrmse_gt_minus_plain"] = learned["final_nrmse_vs_gt_aligned"] - plain["final_nrmse_vs_gt_aligned"]
        line["learned_final_vs_plain_final_nrmse"] = (torch.linalg.vector_norm(xl - xp).item()
                                                      / torch.linalg.vector_norm(xp).item())
        line["learned_vs_plain_time_ratio"] = il["solve_s"] / ip["solve_s"]
        if learned["time_to_1e-3_vs_ref_s"] is not None and plain["time_to_1e-3_vs_ref_s"]:
            line["learned_vs_plain_time_to_1e-3_ratio"] = (learned["time_to_1e-3_vs_ref_s"]
                                                           / plain["time_to_1e-3_vs_ref_s"])
    return line


def bench_lm(args):
    """LMsolver at 256x192x256, 8 coils, 64 shots of 768 lines, readout ROI
    'H' (planes 85..255, solveXT.m:345): the partial derivatives
    (dis_lm_partials: transform and its gradient's DUAL passes, seven coil
    encodes of the ROI, the 28 FP64 sums per state) and one update
    (dis_lm: partials, FP64 6x6 solves, trial transforms and energies),
    plain and with fractional finite-difference weights of order 1 (E.Fd:
    the readout DFT and weighting of the residual and the six Jacobians).
    The sums read Yr, y and the six Jacobians over the ROI once:
    8 Ns nF C 8 bytes per call."""
    import numpy as np
    import torch
    import disorder_amd as D
    from disorder_amd import pyramid as PY
    from disorder_amd import synth

    dev = torch.device("cuda", 0)
    n, C, S = (256, 192, 256), 8, args.shots
    t = int(round(math.sqrt(S)))
    _, k1, k2, off = synth.disorder_sampling(n[0], n[1], S, t, S // t, seed=1)
    x = synth.phantom(n, 40, seed=2, device=dev)
    Sm = synth.coil_maps(n, C, seed=3, device=dev)
    T_gt = synth.random_walk_motion(S, seed=4)
    ctx = D.Context(0)
    E = D.Encoding(Sm, k1, k2, off, ctx=ctx)
    y = E.encode(x, T_gt)
    rng = np.random.default_rng(5)
    T0 = T_gt + rng.normal(0.0, 1.0, (S, 6)) * np.array([0.3, 0.3, 0.3, 0.005, 0.005, 0.005])
    nF = PY.readout_roi(n[2], "H")
    E.set_readout_roi(nF)
    Ns, L = int(off[-1]), nF[1] - nF[0]

    def partials(reps):
        D.lm_partials(y, E, x, T0)
        torch.cuda.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            D.lm_partials(y, E, x, T0)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / reps
        prof = ctx.profile_read()
        ctx.profile(False)
        nl, kms, kb = prof["pass"]
        return ms, {"launches": nl / reps, "ms": kms / reps, "algorithmic_GBps": kb / (kms * 1e6) if kms else None}

    def update():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, cT, _, ntr = D.LMsolver(y, E, x, T0, np.full(S, 1e-3), np.zeros(S, bool))
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0, int(ntr), int(np.sum(cT))

    out = {}
    for label, order in (("plain", 0.0), ("fractional_order_1", 1.0)):
        E.set_lm_filter(PY.fractional_weights(n, L, (1, 1, 1), order, k1, k2) if order else None)
        ms_p, dual = partials(max(1, args.steps // 2))
        ms_u, ntr, nconv = update()
        out[label] = {"lm_partials_ms": ms_p, "dual_passes": dual, "lm_update_ms": ms_u, "trial_rounds": ntr,
                      "converged_states": nconv}
    E.close()
    return {"metric": "LMsolver partial derivatives per call (64 states), ms", "value": out["plain"]["lm_partials_ms"],
            "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "complex64",
            "data": "synthetic (seeded phantom, 8 Gaussian coil maps, DISORDER 8x8 sampling, random-walk motion, "
                    "estimate 0.3 vox / 0.3 deg off)",
            "config": {"workload": f"LMsolver at {n[0]}x{n[1]}x{n[2]}, {C} coils, {S} shots, readout ROI 'H' "
                                   f"({L} of {n[2]} planes)"},
            "lm_reduce_bytes_per_call": 8 * Ns * L * C * 8, **out}


def bench_compress(args):
    """Coil compression at 256x192x256 with 64 coils -> 8 (compressCoils):
    Gram matrix, host eigenbasis, and the split-bf16 MFMA contraction of the
    maps [12.58M x 64] and of the compact data [49152 x 256 x 64].  Errors
    of the compressed arrays against an FP64 product on a voxel sample."""
    import numpy as np
    import torch
    import disorder_amd as D
    from disorder_amd import synth

    dev = torch.device("cuda", 0)
    n, C, Cc, S = (256, 192, 256), 64, 8, args.shots
    Sm = synth.coil_maps(n, C, seed=3, device=dev)
    g = torch.Generator(device=dev).manual_seed(7)
    y = D.empty_colmajor((n[0] * n[1], n[2], C), dev)
    y.copy_(torch.randn(tuple(y.shape), dtype=torch.complex64, device=dev, generator=g))
    ctx = D.default_context(0)
    P = D.coil_gram(Sm, ctx)
    U, Dv, cc = D.coil_basis(P, Cc)
    A = np.conj(U[:, :Cc]).T
    res = {}
    for terms in (3, 6):
        for _ in range(args.warmup):
            D.coil_apply(A, Sm, terms, ctx)
        ctx.profile(True)
        for _ in range(args.steps):
            out = D.coil_apply(A, Sm, terms, ctx)
        prof = ctx.profile_read()
        ctx.profile(False)
        nl, kms, kb = prof["coil_apply"]
        # FP64 check on 4096 sampled voxels
        rng = np.random.default_rng(0)
        idx = torch.from_numpy(rng.choice(Sm.shape[0] * Sm.shape[1] * Sm.shape[2], 4096, replace=False)).to(dev)
        Xs = Sm.reshape(-1, C)[idx].cpu().numpy().astype(np.complex128)
        Os = out.reshape(-1, Cc)[idx].cpu().numpy().astype(np.complex128)
        ref = Xs @ A.T
        err = float(np.linalg.norm(Os - ref) / np.linalg.norm(ref))
        res[terms] = {"avg_launch_ms": kms / nl, "GBps": kb / (kms * 1e6), "frac": kb / (kms * 1e6) / HBM_PEAK_GBS,
                      "rel_l2_vs_fp64": err}
        del out
    ctx.profile(True)
    for _ in range(max(1, args.steps // 2)):
        D.coil_gram(Sm, ctx)
    prof = ctx.profile_read()
    ctx.profile(False)
    gl, gms, _ = prof["coil_gram"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Sc, yc, cc2, _, _ = D.compressCoils(Sm, Cc, y, 3, ctx)
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t0
    gram_flop = 8.0 * Sm.numel() * C
    line = {"metric": "coil compression 64->8 at 256x192x256 (apply GB/s)", "value": res[3]["GBps"], "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "dtype": "bf16x3 MFMA, fp32 accumulate", "data": "synthetic (64 Gaussian coil maps, random data)",
            "config": {"workload": "compressCoils: Gram [12.58M x 64]^H [12.58M x 64], apply [12.58M x 64].[64 x 8] "
                                   "on maps and [12.58M x 64].[64 x 8] on data"},
            "roofline": {"bound": "hbm", "kernel": "coil_apply", "achieved": res[3]["GBps"], "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": res[3]["frac"], "traffic": None},
            "apply": res, "gram_ms": gms / gl, "gram_TFLOPs": gram_flop / (gms / gl * 1e9),
            "compressCoils_total_s": t_all, "components": int(cc2[0])}
    return line


def bench_batched(args):
    """Six independent reconstructions of uneven sizes (160x120x160,
    2x 192x144x192, 224x160x224, 2x 256x192x256; 40-96 shots with 10 %
    truncated; 8 coils) solved by one batched CG of 12 iterations whose
    iterations 2..12 replay one captured hipGraph.  The autotuner picks the
    pencil-pass shapes of each grid first.  Reports the batched time per
    iteration (graph and eager replay), aggregate throughput and each
    volume's latency when solved alone."""
    import numpy as np
    import torch
    import disorder_amd as D
    from disorder_amd import synth

    dev = torch.device("cuda", 0)
    nIt = 12
    P = synth.batched_problem(dev, C=8, seed=11)
    ctx = D.default_context(0)
    t0 = time.perf_counter()
    tune = {}
    for p in P:
        key = "x".join(str(v) for v in p["n"])
        if key not in tune:
            tune[key] = D.autotune(p["n"], S=16, ctx=ctx)
    tune_s = time.perf_counter() - t0
    Es = [p["E"] for p in P]
    ys = [p["y"] for p in P]
    Ts = [p["T"] for p in P]
    Ses = [p["Se"] for p in P]
    D.CGsolver_batched(ys, Es, Ts, Ses, nIt=2, use_graph=True)  # warm-up
    runs = {}
    for mode in ("graph", "eager"):
        best = None
        for _ in range(max(1, args.steps // 5)):
            torch.cuda.synchronize()
            tw = time.perf_counter()
            xs, res, ms = D.CGsolver_batched(ys, Es, Ts, Ses, nIt=nIt, use_graph=(mode == "graph"))
            torch.cuda.synchronize()
            wall = time.perf_counter() - tw
            if best is None or wall < best[0]:
                best = (wall, ms, res)
        wall, ms, res = best
        runs[mode] = {"wall_s": wall, "setup_ms": ms[0], "first_iteration_ms": ms[1],
                      "replayed_iterations_ms": ms[2], "ms_per_iteration_replayed": ms[2] / (nIt - 1),
                      "final_relres": [float(r[-1]) for r in res]}
    # per-volume latency alone (same batched entry point, one volume)
    per_vol = []
    for p in P:
        torch.cuda.synchronize()
        tw = time.perf_counter()
        _, _, ms = D.CGsolver_batched([p["y"]], [p["E"]], [p["T"]], [p["Se"]], nIt=nIt, use_graph=True)
        torch.cuda.synchronize()
        per_vol.append({"n": list(p["n"]), "shots": p["S"], "lines": [int(p["lines"].min()), int(p["lines"].max())],
                        "latency_s": time.perf_counter() - tw, "ms_per_iteration": ms[2] / (nIt - 1)})
    vox = sum(int(np.prod(p["n"])) for p in P)
    g = runs["graph"]
    line = {"metric": "batched CG iterations/s over six uneven volumes", "value": 1000.0 / g["ms_per_iteration_replayed"],
            "unit": "batched iterations/s", "n_gpus": 1, "steps": nIt, "higher_is_better": True,
            "dtype": "complex64 (FP64 CG scalars)", "data": "synthetic (seeded phantoms, 8 coils, truncated DISORDER shots)",
            "config": {"workload": "batched CG, 12 iterations, 6 volumes: " +
                                   ", ".join(f"{p['n'][0]}x{p['n'][1]}x{p['n'][2]}/{p['S']} shots" for p in P),
                       "total_voxels": vox},
            "aggregate_volume_iterations_per_s": 6 * 1000.0 / g["ms_per_iteration_replayed"],
            "aggregate_Gvoxel_iterations_per_s": vox / (g["ms_per_iteration_replayed"] * 1e6),
            "graph": runs["graph"], "eager": runs["eager"], "per_volume": per_vol, "autotune": tune,
            "autotune_s": tune_s}
    return line


def bench_cs(args):
    """Shearlet regulariser of the CS final reconstruction at 256x192x256
    (CSsolver.m / regularize.m:27-50): 50 bands, each IFFT3 -> IRLS weight ->
    FFT3 with the band evaluated from its two plane filters inside the
    passes; 306 FFT-axis passes per application.  Also the IRLS weights and
    one regulariser trace probe."""
    import numpy as np
    import torch
    from disorder_amd import shearlet as SH
    from disorder_amd import synth

    dev = torch.device("cuda", 0)
    n = (256, 192, 256)
    x = synth.phantom(n, 12, seed=8, device=dev)
    t0 = time.perf_counter()
    sysh = SH.ShearletSystem(n)
    build_s = time.perf_counter() - t0
    ctx = sysh.ctx
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sysh.weights(x)
    torch.cuda.synchronize()
    weights_s = time.perf_counter() - t0
    out = torch.zeros_like(x)
    for _ in range(args.warmup):
        sysh.regularize(x, 0.5, out)
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sysh.regularize(x, 0.5, out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    nl, kms, kb = prof["shearlet_pass"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = sysh.energy(x)
    torch.cuda.synchronize()
    energy_s = time.perf_counter() - t0
    sysh.close()
    ms = wall / args.steps * 1e3
    line = {"metric": "shearlet regulariser applications/s at 256x192x256 (50 bands)", "value": 1e3 / ms,
            "unit": "apps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "complex64",
            "data": "synthetic (seeded phantom)",
            "config": {"workload": "CSsolver shearlet term: IFFT3(sum_w S_w FFT3(We_w IFFT3(conj(S_w) FFT3 x))) at "
                                   "256x192x256, buildShearlet(N,1,.,1,'kos'), 50 bands"},
            "roofline": {"bound": "hbm", "kernel": "shearlet_pass", "achieved": kb / (kms * 1e6), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": kb / (kms * 1e6) / HBM_PEAK_GBS, "traffic": None,
                         "algorithmic_bytes_per_launch": kb / nl, "avg_launch_ms": kms / nl,
                         "launches_per_step": nl / args.steps},
            "weights_s": weights_s, "energy_probe_s": energy_s, "host_plane_build_s": build_s,
            "energy_check": e}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = _cpu_baseline_cs(n)
    return line


def _cpu_baseline_cs(n):
    """FP64 oracle (scipy.fft, 16 threads) on one of the 50 bands at the full
    grid: t_band; apps/s = 1 / (50 t_band + t_fft3 pair)."""
    import numpy as np
    import scipy.fft as sf
    from oracle import shearlet_oracle as SO
    n1, n2, n3 = n
    A = SO.shearlets_2d(n1, n2)[:, :, 2]
    B = np.transpose(SO.shearlets_2d(n3, n2), (1, 0, 2))[:, :, 2]
    S = A[:, :, None] * B[None, :, :]
    rng = np.random.default_rng(0)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    We = rng.uniform(0.5, 2.0, n)
    t0 = time.perf_counter()
    xF = sf.fftn(x, workers=16)
    t1 = time.perf_counter()
    t = sf.ifftn(np.conj(S) * xF, workers=16) * We
    acc = sf.fftn(t, workers=16) * S
    t2 = time.perf_counter()
    tb, tf = t2 - t1, t1 - t0
    return {"value": 1.0 / (50 * tb + 2 * tf), "unit": "apps/s", "cores": 16, "kind": "port",
            "sample": f"FP64 oracle band body (IFFT3, weight, FFT3, band multiply) of 1 of 50 bands at "
                      f"256x192x256: {tb:.2f} s, FFT3 {tf:.2f} s; apps/s = 1/(50 band + 2 FFT3)"}


def bench_proposal(args):
    """Acceptance of the learned proposal against the distance of the
    current estimate from the solution (learned.acceptance_curve): 3
    simulated 64^3 scans of 32 shots and 8 coils (the training
    distribution), per-shot errors N(0, d) for d from 0.05 to 4 voxels /
    degrees, without and with a common offset N(0, 2) shared by the shots,
    image from 4 CG iterations; the proposal (Newton step plus the
    network's correction, centred) beside the diagonal Newton step alone.
    Rows are averaged over the scans."""
    import numpy as np
    from disorder_amd import learned as LN
    wpath = args.learned if args.learned not in (None, "", "none") else LN.WEIGHTS
    net = LN.MotionNet.load(wpath)
    dists = (0.05, 0.1, 0.2, 0.5, 1.0, 2.0, 4.0)
    offsets = (0.0, 2.0)
    t0 = time.perf_counter()
    rows = LN.acceptance_curve(net, np.random.default_rng(11), dists, offsets=offsets, n_scans=3)
    out = []
    for o in offsets:
        for d in dists:
            rs = [r for r in rows if r["dist"] == d and r["offset"] == o]
            row = {"dist": d, "offset": o, "err_before": float(np.mean([r["err_before"] for r in rs]))}
            for k in ("net", "newton"):
                row[k] = {m: float(np.mean([r[k][m] for r in rs])) for m in ("accept", "err_all", "err_accepted")}
            out.append(row)
    return {"metric": "learned-proposal acceptance against distance from the solution", "unit": "fraction of shots",
            "n_gpus": 1, "data": "synthetic (learned.simulate_scan: 64^3, 8 coils, 32 shots, 3 scans)",
            "config": {"workload": "per-shot errors N(0, d) voxel/degree, common offset N(0, o), image from 4 CG "
                                   "iterations; err_* = rms per-shot error with the common motion removed"},
            "weights": os.path.relpath(wpath, os.path.dirname(os.path.abspath(__file__))),
            "elapsed_s": time.perf_counter() - t0, "rows": out}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (sharding the shots over GPUs "
              "is designed, not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "pyramid":
        print(json.dumps(bench_pyramid(args)), flush=True)
        return
    if args.workload == "compress":
        print(json.dumps(bench_compress(args)), flush=True)
        return
    if args.workload == "cs":
        print(json.dumps(bench_cs(args)), flush=True)
        return 0
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "small":
        print(json.dumps(bench_small(args)), flush=True)
        return
    if args.workload == "fold":
        print(json.dumps(bench_fold(args)), flush=True)
        return
    if args.workload == "stream":
        print(json.dumps(bench_stream(args)), flush=True)
        return
    if args.workload == "proposal":
        print(json.dumps(bench_proposal(args)), flush=True)
        return
    if args.workload == "lm":
        print(json.dumps(bench_lm(args)), flush=True)
        return
    import numpy as np
    import torch
    import disorder_amd as D
    from disorder_amd import synth

    dev = torch.device("cuda", 0)
    n = (256, 192, 256)
    C = 8
    S = args.shots
    t = int(round(math.sqrt(S)))
    _, k1, k2, off = synth.disorder_sampling(n[0], n[1], S, t, S // t, seed=1)
    x = synth.phantom(n, 40, seed=2, device=dev)
    Sm = synth.coil_maps(n, C, seed=3, device=dev)
    T = synth.random_walk_motion(S, seed=4)
    ctx = D.Context(0)
    E = D.Encoding(Sm, k1, k2, off, ctx=ctx)
    out = D.empty_colmajor(n, dev)
    for _ in range(args.warmup):
        E.encodedecode(x, T, out=out)
    torch.cuda.synchronize()
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E.encodedecode(x, T, out=out)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    ms = dt * 1000.0 / args.steps
    # apps/s as 1 / median of 20 single applications, each synchronised (after the timed steps)
    single = []
    for _ in range(20):
        t1 = time.perf_counter()
        E.encodedecode(x, T, out=out)
        torch.cuda.synchronize()
        single.append((time.perf_counter() - t1) * 1000.0)
    med = float(np.median(single))
    kernels = {}
    tf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "ehe_traffic.json")
    measured = {}
    if os.path.exists(tf) and S == 64:
        with open(tf) as fh:
            measured = json.load(fh)["bytes_per_launch"]
    for name, (nl, kms, kb) in prof.items():
        if nl:
            k = {"launches_per_step": nl / args.steps, "ms_per_step": kms / args.steps, "avg_launch_ms": kms / nl}
            if name == "coil_stage":
                # the pass model counts the per-coil hybrid volumes as HBM traffic, but the coil
                # stage keeps them on chip: bytes / time is an equivalent, not a transfer rate
                k["pass_model_equivalent_GBps"] = kb / (kms * 1e6) if kms else None
            else:
                k["algorithmic_GBps"] = kb / (kms * 1e6) if kms else None
            if name in measured and kms:
                # the file holds the coil stage per E^H E (its two launches), the passes per launch
                per = kms / args.steps if name == "coil_stage" else kms / nl
                k["counter_traffic_GBps"] = measured[name] / (per * 1e6)
            kernels[name] = k
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    nl, kms, kb = prof[dom]
    achieved = kb / (kms * 1e6)
    V = 8 * n[0] * n[1] * n[2]
    passmodel = S * V * (4 * 10 + 6 * C + 3)
    traffic = measured.get(dom)
    line = {
        "metric": "E^H E apps/s at 256x256x192x8 coils; time to 1e-3 NRMSE, learned vs plain LM",
        "value": 1000.0 / ms,
        "unit": "E^H E apps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic (seeded phantom, 8 Gaussian coil maps, DISORDER 8x8 random-checkered sampling, "
                "random-walk motion)",
        "config": {"workload": f"E^H E (dis_encodedecode) at 256x192x256 (PE1 x PE2 x RO), {C} coils, {S} shots "
                               f"of {off[1] - off[0]} lines", "grid": list(n), "coils": C, "shots": S},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/ehe_traffic.json (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1: L2 fills "
                                       "and write-backs, served by the MALL or HBM)",
                     "algorithmic_bytes_per_launch": kb / nl, "avg_launch_ms": kms / nl},
        "pass_model_v1": {"bytes_per_app": passmodel, "GBps": passmodel / (ms * 1e6),
                          "frac_of_peak": passmodel / (ms * 1e6) / HBM_PEAK_GBS},
        "kernels": kernels,
        "median_of_20": {"ms": med, "apps_per_s": 1000.0 / med, "min_ms": min(single), "max_ms": max(single)},
    }
    cp = stream_copy_rate(dev)
    line["stream_copy"] = cp
    line["pass_model_v1"]["frac_of_stream_copy"] = line["pass_model_v1"]["GBps"] / cp["GBps"]
    line["roofline"]["frac_of_stream_copy"] = achieved / cp["GBps"]
    E.close()
    del E, x, Sm, out
    torch.cuda.empty_cache()
    if not args.no_pyramid:
        # the metric's second half: time to 1e-3 NRMSE of the joint pyramid, plain and learned
        # (bench.py --workload pyramid gives the full record)
        try:
            pl = bench_pyramid(args)
            lr = pl.get("learned", {})
            line["time_to_1e-3_nrmse"] = {
                "plain_s": pl["plain"]["time_to_1e-3_vs_ref_s"], "learned_s": lr.get("time_to_1e-3_vs_ref_s"),
                "learned_vs_plain_time_to_1e-3_ratio": pl.get("learned_vs_plain_time_to_1e-3_ratio"),
                "plain_closing_cg": {k: pl["plain"]["crossing"].get(k) for k in
                                     ("m", "closing_s_to_m", "closing_iterations", "closing_s", "t_before_s")},
                "learned_closing_cg": {k: lr.get("crossing", {}).get(k) for k in
                                       ("m", "closing_s_to_m", "closing_iterations", "closing_s", "t_before_s",
                                        "nrmse_vs_ref_at")},
                "learned_final_vs_plain_final_nrmse": pl.get("learned_final_vs_plain_final_nrmse"),
                "learned_vs_plain_time_ratio": pl.get("learned_vs_plain_time_ratio"),
                "plain_solve_s": pl["plain"]["solve_s"],
                "plain_final_nrmse_vs_ref": pl["plain"]["final_nrmse_vs_ref"],
                "plain_final_nrmse_vs_gt_aligned": pl["plain"]["final_nrmse_vs_gt_aligned"],
                "known_motion_cg_nrmse_vs_gt": pl["known_motion_cg_nrmse_vs_gt"],
                "plain_motion_err_max": [pl["plain"]["motion_err_max_vox"], pl["plain"]["motion_err_max_deg"]],
                "learned_nrmse_gt_minus_plain": pl.get("learned_nrmse_gt_minus_plain"),
                "learned_proposals": lr.get("proposals"),
                "reference": pl["config"]["reference"],
                "workload": pl["config"]["workload"], "unit": "s", "higher_is_better": False}
        except Exception as exc:  # reported, not fatal to the E^H E line
            line["time_to_1e-3_nrmse"] = {"error": repr(exc)}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, C, S, seed=5)
    print(json.dumps(line), flush=True)


def bench_small(args):
    """The reference's CPU-runnable case at 64^3: 10 random ellipsoids
    (magnitudes U(0.2, 1), quadratic phase within +-pi/4), 8 Gaussian coils
    (sigma 24, radius 48 voxels), 16 shots of 4x4 random-checkered tiles
    (256 lines each), rotations U(-5, 5) deg and translations U(-3, 3)
    voxels, 30 dB noise; 20 CG iterations with the known motion, then the
    joint CG+LM pyramid over 3 levels.  value = GPU wall time of both steps.
    cpu_baseline = the FP64 oracle on the same inputs: by default a bounded
    sample (3 of the 20 known-motion CG iterations, scaled to 20, to compare
    with cg20_s; the pyramid is not sampled); with --cpu-full the whole run,
    and the distance of the GPU's image and motion to the oracle's."""
    import numpy as np
    import torch
    import disorder_amd as D
    from disorder_amd import pyramid as PY
    from disorder_amd import synth

    dev = torch.device("cuda", 0)
    n, C, S = (64, 64, 64), 8, 16
    _, k1, k2, off = synth.disorder_sampling(n[0], n[1], S, 4, 4, seed=1)
    x_gt = synth.phantom(n, 10, seed=2, device=dev, mag=(0.2, 1.0))
    Sm = synth.coil_maps(n, C, seed=3, device=dev)
    T_gt = synth.uniform_motion(S, seed=4)
    E = D.Encoding(Sm, k1, k2, off)
    y = E.encode(x_gt, T_gt)
    y = synth.add_noise(y, synth.noise_level(x_gt, Sm, 30.0), seed=5)
    kw = dict(nlevels=3, n_outer=args.outer, n_cg=args.cg, final_cg=args.final_cg)
    for _ in range(args.warmup):  # first-call set-up (plans, twiddles) outside the timed runs
        D.CGsolver(y, E, T_gt, nIt=2, tolType="RelativeResidual2Error", tol=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xk, _, _ = D.CGsolver(y, E, T_gt, nIt=20, tolType="RelativeResidual2Error", tol=0)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    E.close()
    t0 = time.perf_counter()
    x, T, info = PY.solveXT(y, Sm, k1, k2, off, n, **kw)
    torch.cuda.synchronize()
    t_pyr = time.perf_counter() - t0
    ng = torch.linalg.vector_norm(x_gt).item()
    T_err = T_gt - T
    T_err = T_err - T_err.mean(axis=0, keepdims=True)
    # the joint estimate is fixed up to a motion common to all shots: move it
    # by mean(T - T_gt) before comparing with the ground truth
    xa = D.sincRigidTransform(x, np.mean(T - T_gt, axis=0, keepdims=True), di=1)[..., 0]
    line = {"metric": "64^3 reference case: known-motion CG (20 it) + 3-level joint pyramid, wall s",
            "value": t_cg + t_pyr, "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "complex64",
            "data": "synthetic (seeded: 10 ellipsoids, 8 Gaussian coils, 4x4 random-checkered sampling, "
                    "uniform motion 5 deg / 3 voxels, 30 dB noise)",
            "config": {"workload": f"64x64x64, {C} coils, {S} shots of {off[1] - off[0]} lines; CG 20 it "
                                   f"(known motion), then pyramid 3 levels, {args.outer} outer, {args.cg} CG, "
                                   f"final CG {args.final_cg}"},
            "cg20_s": t_cg, "pyramid_s": t_pyr, "levels": info["levels"],
            "cg_nrmse_vs_gt": torch.linalg.vector_norm(xk - x_gt).item() / ng,
            "pyramid_nrmse_vs_gt": torch.linalg.vector_norm(x - x_gt).item() / ng,
            "pyramid_nrmse_vs_gt_aligned": torch.linalg.vector_norm(xa - x_gt).item() / ng,
            "motion_err_max_vox": float(np.abs(T_err[:, :3]).max()),
            "motion_err_max_deg": float(np.abs(T_err[:, 3:]).max() * 180 / math.pi)}
    if args.no_cpu_baseline:
        return line
    from oracle import disorder_oracle as O
    workers = min(os.cpu_count() or 1, 16)
    O.set_fft_workers(workers)
    yh = np.asarray(y.cpu().numpy(), np.complex128)
    Sh = np.asarray(Sm.cpu().numpy(), np.complex128)
    samp = O.Sampling(k1, k2, off, n[0], n[1])
    kcg = 20 if args.cpu_full else 3
    t0 = time.perf_counter()
    O.cg_solver(yh, Sh, samp, T_gt, n, nIt=kcg, tolType="RelativeResidual2Error", tol=0)
    t_ocg = (time.perf_counter() - t0) * 20 / kcg
    print(f"[small] oracle CG {kcg} it: {t_ocg * kcg / 20:.1f} s", file=sys.stderr, flush=True)
    cpu = {"unit": "s", "cores": workers, "kind": "port", "cg20_s": t_ocg}
    if args.cpu_full:
        def cb(level, it, xo, To):
            print(f"[small] oracle pyramid level {level} outer {it}: {time.perf_counter() - t1:.1f} s",
                  file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        xo, To = O.pyramid_solve(yh, Sh, k1, k2, off, n, callback=cb, **kw)
        t_opyr = time.perf_counter() - t1
        xg = x.cpu().numpy()
        cpu.update({"value": t_ocg + t_opyr, "pyramid_s": t_opyr,
                    "sample": f"the whole run: FP64 oracle (scipy.fft, {workers} threads) CG 20 it + pyramid"})
        line["parity_vs_oracle"] = {
            "image_nrmse": float(np.linalg.norm(xg - xo) / np.linalg.norm(xo)),
            "motion_max_vox": float(np.abs(T[:, :3] - To[:, :3]).max()),
            "motion_max_deg": float(np.abs(T[:, 3:] - To[:, 3:]).max() * 180 / math.pi)}
    else:
        cpu.update({"value": t_ocg, "compare_to": "cg20_s",
                    "sample": f"FP64 oracle (scipy.fft, {workers} threads) known-motion CG, {kcg} of the 20 "
                              f"iterations timed and scaled to 20; the pyramid is timed only with --cpu-full"})
    O.set_fft_workers(None)
    line["cpu_baseline"] = cpu
    return line


def bench_fold(args):
    """E^H E with SENSE folding of the phase encodes: maps on 256x192x256,
    data acquired on a 256x128 phase-encode grid (factor 1.5 along the second
    phase encode), 8 coils, 64 shots of 8x8 random-checkered tiles on the
    acquired grid; the same timing and per-kernel split as the default line."""
    import torch
    import disorder_amd as D
    from disorder_amd import synth

    dev = torch.device("cuda", 0)
    n, NY, C, S = (256, 192, 256), (256, 128), 8, args.shots
    _, k1, k2, off = synth.disorder_sampling(NY[0], NY[1], S, 8, 8, seed=1)
    x = synth.phantom(n, 40, seed=2, device=dev)
    Sm = synth.coil_maps(n, C, seed=3, device=dev)
    T = synth.random_walk_motion(S, seed=4)
    ctx = D.Context(0)
    E = D.Encoding(Sm, k1, k2, off, ctx=ctx, NY=NY)
    out = D.empty_colmajor(n, dev)
    for _ in range(args.warmup):
        E.encodedecode(x, T, out=out)
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E.encodedecode(x, T, out=out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    prof = ctx.profile_read()
    ctx.profile(False)
    E.close()
    kernels = {k: {"launches_per_step": nl / args.steps, "ms_per_step": kms / args.steps,
                   "avg_launch_ms": kms / nl} for k, (nl, kms, kb) in prof.items() if nl}
    return {"metric": "E^H E apps/s with SENSE folding", "value": 1000.0 / ms, "unit": "E^H E apps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "dtype": "complex64", "data": "synthetic (seeded phantom, 8 Gaussian coil maps, random-checkered "
                                          "sampling on the acquired grid, random-walk motion)",
            "config": {"workload": f"E^H E with folding: maps {n[0]}x{n[1]}x{n[2]}, acquired {NY[0]}x{NY[1]}, "
                                   f"{C} coils, {S} shots of {off[1] - off[0]} lines"},
            "kernels": kernels}


def bench_stream(args, nbytes=2 << 30):
    """The streaming ceiling: dis_stream_probe's copy kernels (several access
    shapes, block counts) over 2 GiB, read + write counted, plus torch's
    copy_.  value = the best rate."""
    import ctypes
    import torch
    import disorder_amd as D
    from disorder_amd import _lib
    dev = torch.device("cuda", 0)
    ctx = D.Context(0)
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    rows = []
    names = {0: "grid-stride 8 B lanes", 1: "grid-stride 16 B lanes", 2: "grid-stride 4x16 B in flight",
             3: "grid-stride 4x16 B nontemporal", 4: "pass tile (32 KiB per workgroup)",
             5: "pass tile, next tile's loads before the stores",
             6: "readout-axis pass pattern (256x192x256 volumes)", 7: "dim-2 pass pattern (256x192x256 volumes)"}
    vol = 256 * 192 * 256 * 8
    for v, blocks in [(0, ncu * 8), (1, ncu * 8), (2, ncu * 4), (2, ncu * 8), (2, ncu * 16), (3, ncu * 8),
                      (3, ncu * 16), (4, 0), (5, ncu * 4), (5, ncu * 8), (5, ncu * 16), (6, 0), (7, 0)]:
        ms = ctypes.c_double(0)
        _lib.check(ctx.lib.dis_stream_probe(ctx.h, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                            nbytes, v, blocks, max(args.steps, 5), ctypes.byref(ms)), ctx.h)
        moved = 2.0 * (nbytes // vol) * vol if v >= 6 else 2.0 * nbytes
        rows.append({"variant": v, "shape": names[v], "blocks": blocks, "ms": ms.value,
                     "GBps": moved / (ms.value * 1e6)})
        print(f"[stream] {names[v]}, {blocks} blocks: {rows[-1]['GBps']:.0f} GB/s", file=sys.stderr, flush=True)
    tc = stream_copy_rate(dev, nbytes)
    best = max((r for r in rows if r["variant"] < 6), key=lambda r: r["GBps"])
    return {"metric": "streaming copy rate (read + write)", "value": best["GBps"], "unit": "GB/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "bytes", "data": "2 GiB device buffers",
            "config": {"workload": "device-to-device copy of 2 GiB, best of the flat probe shapes (the two pass "
                                   "patterns are reported beside them)"},
            "best": best, "frac_of_peak": best["GBps"] / HBM_PEAK_GBS, "probes": rows, "torch_copy": tc}


def stream_copy_rate(dev, nbytes=2 << 30, reps=10):
    """Achievable HBM rate on this GPU: a device-to-device copy of nbytes
    (read + write counted), best of reps, timed with events on torch's stream."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        t = e0.elapsed_time(e1)
        best = t if best is None else min(best, t)
    del a, b
    return {"GBps": 2.0 * nbytes / (best * 1e6), "bytes_moved": 2 * nbytes, "ms": best,
            "how": "torch copy_ of a 2 GiB float32 buffer, best of 10, read + write counted"}


if __name__ == "__main__":
    main()
