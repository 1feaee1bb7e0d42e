"""Benchmark: E^H E applications per second on one MI355X.

Workload: the encoding operator of DISORDER's reconstruction path at
256x192x256 (phase encode 256x192, readout 256), 8 coils, 64 DISORDER shots
(8x8 random-checkered tiles, 768 phase-encode lines per shot), random-walk
rigid motion per shot, FP32 complex.  One step = one E^H We E x
(dis_encodedecode) with every input already resident in HBM.

Prints one JSON line.  ``roofline`` is measured live with HIP events recorded
by the library around every launch of the dominant kernel during the timed
steps; ``cpu_baseline`` times the FP64 CPU oracle on a bounded sample (one
shot's E^H E contribution) on this host.
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
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--workload", choices=["ehe", "pyramid", "compress"], default="ehe",
                    help="ehe: E^H E apps/s (default); pyramid: time to 1e-3 NRMSE of the joint CG+LM pyramid; "
                         "compress: 64 -> 8 coil compression of maps and data")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--outer", type=int, default=10)
    ap.add_argument("--cg", type=int, default=8)
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
    return {"value": 1.0 / (S * dt), "unit": "E^H E apps/s", "cores": workers, "kind": "port",
            "sample": f"FP64 oracle (scipy.fft, {workers} threads) E^H E of 1 of {S} shots at "
                      f"{n[0]}x{n[1]}x{n[2]}, {C} coils: {dt:.2f} s; apps/s = 1/({S} x that)"}


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


def bench_pyramid(args):
    """Time to 1e-3 NRMSE of the plain joint reconstruction (4 levels, 10 LM
    iterations per level, 8 CG iterations per outer step).  x_ref is the
    method's own converged image; NRMSE_t = ||x_t - x_ref|| / ||x_ref|| of
    the image after each outer iteration, upsampled to the full grid (the
    upsampling and snapshot copies are outside the timed solve)."""
    import numpy as np
    import torch
    import disorder_amd as D
    from disorder_amd import pyramid as PY

    dev = torch.device("cuda", 0)
    torch.cuda.synchronize()
    tg = time.perf_counter()
    P = reconstruction_problem(dev, S=args.shots)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - tg
    n = P["n"]
    snaps = []

    def cb(level, it, x, T, t_solve):
        xf = D.resampling(x, n) if list(x.shape) != list(n) else x.clone()
        snaps.append((level, it, t_solve, xf))
        print(f"[pyramid] level {level} outer {it}: solve {t_solve:.2f} s", file=sys.stderr, flush=True)

    x, T, info = PY.solveXT(P["y"], P["S"], P["k1"], P["k2"], P["off"], n, nlevels=args.levels,
                            n_outer=args.outer, n_cg=args.cg, callback=cb)
    nr = torch.linalg.vector_norm(x).item()
    xg = P["x_gt"]
    curve = []
    tts = None
    for lv, it, ts, xs in snaps:
        e = torch.linalg.vector_norm(xs - x).item() / nr
        curve.append({"level": lv, "outer": it, "t_s": ts, "nrmse_vs_final": e})
        if tts is None and e <= 1e-3:
            tts = ts
    del snaps
    T_err = P["T_gt"] - T
    T_err = T_err - T_err.mean(axis=0, keepdims=True)
    nrmse_gt = torch.linalg.vector_norm(x - xg).item() / torch.linalg.vector_norm(xg).item()
    line = {"metric": "time to 1e-3 NRMSE (plain LM), s", "value": tts, "unit": "s", "n_gpus": 1,
            "higher_is_better": False, "dtype": "complex64",
            "data": "synthetic (seeded phantom, 32 Gaussian coils compressed to 8, DISORDER 8x8 sampling, "
                    "random-walk motion, 25 dB noise)",
            "config": {"workload": f"joint CG+LM pyramid at {n[0]}x{n[1]}x{n[2]}, 8 coils, {args.shots} shots, "
                                   f"{args.levels} levels, {args.outer} outer, {args.cg} CG"},
            "solve_s": info["solve_s"], "data_gen_s": gen_s, "levels": info["levels"], "lm_trials": info["lm_trials"],
            "nrmse_vs_gt": nrmse_gt,
            "motion_err_max_vox": float(np.abs(T_err[:, :3]).max()),
            "motion_err_max_deg": float(np.abs(T_err[:, 3:]).max() * 180 / math.pi),
            "curve": curve}
    return line


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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (multi-GPU sharding of shots "
              "is a design note in DESIGN.md, not part of this build)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "pyramid":
        print(json.dumps(bench_pyramid(args)), flush=True)
        return
    if args.workload == "compress":
        print(json.dumps(bench_compress(args)), flush=True)
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
    kernels = {}
    for name, (nl, kms, kb) in prof.items():
        if nl:
            kernels[name] = {"launches_per_step": nl / args.steps, "ms_per_step": kms / args.steps,
                             "avg_launch_ms": kms / nl, "algorithmic_GBps": kb / (kms * 1e6) if kms else None}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    nl, kms, kb = prof[dom]
    achieved = kb / (kms * 1e6)
    V = 8 * n[0] * n[1] * n[2]
    passmodel = S * V * (4 * 10 + 6 * C + 3)
    traffic = None
    tf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "ehe_traffic.json")
    if os.path.exists(tf) and S == 64:
        with open(tf) as fh:
            traffic = json.load(fh)["bytes_per_launch"].get(dom)
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
                     "traffic_source": "profiles/ehe_traffic.json (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1)",
                     "algorithmic_bytes_per_launch": kb / nl, "avg_launch_ms": kms / nl},
        "pass_model_v1": {"bytes_per_app": passmodel, "GBps": passmodel / (ms * 1e6),
                          "frac_of_peak": passmodel / (ms * 1e6) / HBM_PEAK_GBS},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, C, S, seed=5)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
