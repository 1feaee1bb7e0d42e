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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (multi-GPU sharding of shots "
              "is a design note in DESIGN.md, not part of this build)", file=sys.stderr)
        sys.exit(2)
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
    line = {
        "metric": "E^H E apps/s at 256x256x192x8 coils",
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
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
