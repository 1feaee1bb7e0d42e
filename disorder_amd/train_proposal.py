"""Train the learned motion proposal on scans made by the method's own
simulator (disorder_amd.learned.simulate_scan; with --pyramid, the states the
plain pyramid visits on those scans, learned.pyramid_examples): phantoms at 64^3 and
128^3, 8 coils, 32-64 DISORDER shots, random-walk motion with step sigma
0.2-1.5 deg / voxel plus 5 % abrupt jumps up to 10 deg / 10 voxels.  Every
scan contributes the features and relative updates (the true update less
its mean over shots) at `--states` current estimates.  Training: Adam, lr
1e-3, batches of up to 512 shot rows of whole scans, L2 loss of the Newton
step plus the network's output to the relative update (learned.net_target);
all on the GPU through libdisorder_hip.so.  The held-out errors compare the
proposal with no update and with the Newton step alone.

The committed weights (disorder_amd/weights/motion_net.npy) come from the
states of 2,000 pyramid solves, kept where the plain solve ends within 1
voxel / 1 degree of the true motion:

  python -m disorder_amd.train_proposal --pyramid --scans 1000 --seed 202 --save-data p2.npz
  python -m disorder_amd.train_proposal --pyramid --scans 1000 --seed 203 --save-data p3.npz
  python -m disorder_amd.train_proposal --load-data p2.npz p3.npz --max-err 1.0 \
      --out disorder_amd/weights/motion_net.npy
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import learned as LN


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scans", type=int, default=20000)
    ap.add_argument("--frac128", type=float, default=0.05, help="fraction of scans simulated at 128^3")
    ap.add_argument("--states", type=int, default=1, help="current estimates per scan")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--val", type=float, default=0.05, help="held-out fraction of the scans")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cg", type=int, default=4, help="CG iterations for the image at each estimate")
    ap.add_argument("--out", default=LN.WEIGHTS)
    ap.add_argument("--log", default=None)
    ap.add_argument("--save-data", default=None, help="write the simulated examples (.npz, float16) and stop")
    ap.add_argument("--load-data", nargs="*", default=None, help="train on saved examples instead of simulating")
    ap.add_argument("--init", default=None, help="start from these weights (fine-tuning)")
    ap.add_argument("--max-err", dest="max_err", type=float, default=None,
                    help="train only on pyramid examples of scans whose final motion error (voxel / degree, common "
                         "motion removed) is at most this")
    ap.add_argument("--pyramid", action="store_true",
                    help="examples from the states the plain pyramid visits (every level and outer iteration) "
                         "instead of estimates drawn around the solution")
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(a.seed)
    data = []
    errs = []  # per example: final motion error of the scan's plain solve (0 for drawn estimates)
    t0 = time.perf_counter()
    if a.load_data:
        for path in a.load_data:
            z = np.load(path, allow_pickle=False)
            F, T, so = z["F"].astype(np.float32), z["T"].astype(np.float32), z["off"]
            er = z["err"] if "err" in z.files else np.zeros(len(so) - 1)
            data += [(F[so[i]:so[i + 1]], T[so[i]:so[i + 1]]) for i in range(len(so) - 1)]
            errs += [float(e) for e in er]
        if a.max_err is not None:
            keep = [i for i, e in enumerate(errs) if e <= a.max_err]
            print(f"[train] {len(keep)} of {len(data)} examples from scans solved within {a.max_err}", flush=True)
            data = [data[i] for i in keep]
            errs = [errs[i] for i in keep]
        print(f"[train] loaded {len(data)} examples", flush=True)
    else:
        n128 = int(round(a.scans * a.frac128))
        failed = 0
        for i in range(a.scans):
            n = (128, 128, 128) if i < n128 else (64, 64, 64)
            S = int(rng.choice([32, 64]))
            if a.pyramid:
                try:
                    rows, err = LN.pyramid_examples(rng, n, S, 8, dev, nlevels=3 if n[0] == 64 else 4)
                    data += rows
                    errs += [err] * len(rows)
                except RuntimeError as exc:  # a solve that ends with an error status: the scan is skipped
                    failed += 1
                    print(f"[train] scan {i}: {exc}", flush=True)
            else:
                E, x, y, T, off, Sm = LN.simulate_scan(rng, n, S, 8, dev)
                rows = LN.training_examples(rng, E, x, y, T, off, n_cg=a.cg, n_states=a.states)
                data += rows
                errs += [0.0] * len(rows)
                E.close()
                del x, y, Sm
            if (i + 1) % (100 if a.pyramid else 500) == 0 or i + 1 == a.scans:
                print(f"[train] simulated {i + 1}/{a.scans} scans ({time.perf_counter() - t0:.0f} s)", flush=True)
    gen_s = time.perf_counter() - t0
    if a.save_data:
        so = np.concatenate([[0], np.cumsum([d[0].shape[0] for d in data])]).astype(np.int64)
        np.savez_compressed(a.save_data, F=np.concatenate([d[0] for d in data]).astype(np.float16),
                            T=np.concatenate([d[1] for d in data]).astype(np.float16), off=so,
                            err=np.asarray(errs, np.float32))
        print(json.dumps({"scans": a.scans, "examples": len(data), "rows": int(so[-1]), "gen_s": gen_s}), flush=True)
        return 0
    order = rng.permutation(len(data))
    nval = int(len(data) * a.val)
    val = [data[i] for i in order[:nval]]
    trn = [data[i] for i in order[nval:]]
    net = LN.MotionNet.load(a.init) if a.init else LN.MotionNet(seed=a.seed)
    t1 = time.perf_counter()

    def log(ep, loss):
        if ep % 5 == 0 or ep == a.epochs - 1:
            print(f"[train] epoch {ep}: loss {loss:.4f} ({time.perf_counter() - t1:.0f} s)", flush=True)
    losses = LN.train(net, trn, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    train_s = time.perf_counter() - t1
    # held-out error against two baselines: no update, and the diagonal Newton step
    Fv = np.concatenate([d[0] for d in val]) if val else np.zeros((0, LN.NIN), np.float32)
    Tv = np.concatenate([d[1] for d in val]) if val else np.zeros((0, LN.NOUT), np.float32)
    so = np.concatenate([[0], np.cumsum([d[0].shape[0] for d in val])]).astype(np.int32)
    res = {"scans": a.scans, "examples": len(data), "rows": int(sum(d[0].shape[0] for d in data)),
           "gen_s": gen_s, "train_s": train_s, "epochs": a.epochs, "loss_first": losses[0], "loss_last": losses[-1]}
    if len(Fv):
        Yv = net.forward(torch.from_numpy(Fv).to(dev), so).cpu().numpy()
        res.update({"val_mse": float(np.mean((Fv[:, :6] + Yv - Tv) ** 2)), "val_mse_zero": float(np.mean(Tv ** 2)),
                    "val_mse_newton": float(np.mean((Fv[:, :6] - Tv) ** 2))})
    net.save(a.out)
    print(json.dumps(res), flush=True)
    if a.log:
        with open(a.log, "w") as fh:
            json.dump({**res, "losses": losses}, fh)
    return 0


if __name__ == "__main__":
    sys.exit(main())
