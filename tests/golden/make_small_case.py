"""Writes tests/golden/small_case_64.npz: the FP64 oracle's run of the
reference's 64^3 case (tests/golden/small_case.py builds the inputs): the
image after 20 CG iterations with the known motion, then the joint pyramid
(3 levels, the nX CG schedule with nIt(2) = 8, the readout ROI of fat-shift
direction 'H', the closing 'Energy' CG) from zero: its image, motion and
the outer iterations run at each level.

    python tests/golden/make_small_case.py      (about 10 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import disorder_oracle as O  # noqa: E402
from tests.golden import small_case as SC  # noqa: E402


def main():
    O.set_fft_workers(os.cpu_count() or 1)
    t0 = time.time()
    d = SC.make_inputs()
    samp = O.Sampling(d["k1"], d["k2"], d["off"], SC.N[0], SC.N[1])
    x_cg, _, nn = O.cg_solver(d["y"], d["Sm"], samp, d["T_gt"], SC.N, nIt=SC.CG_KNOWN,
                              tolType="RelativeResidual2Error", tol=0.0)
    print(f"known-motion CG: {nn} iterations, {time.time() - t0:.0f} s", flush=True)
    outer = {}

    def cb(li, it, x, T):
        outer[li] = it
        print(f"  level {li} outer {it}: {time.time() - t0:.0f} s", flush=True)
    info = {}
    x_pyr, T_pyr = O.pyramid_solve(d["y"], d["Sm"], d["k1"], d["k2"], d["off"], SC.N, callback=cb, info=info,
                                   **SC.PYR)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_case_64.npz")
    np.savez_compressed(out, x_cg=x_cg.astype(np.complex64), x_pyr=x_pyr.astype(np.complex64), T_pyr=T_pyr,
                        outer=np.array([outer[k] for k in sorted(outer)], np.int64), seed=SC.SEED)
    tr = lambda x, T: O.rigid_forward(x, T, SC.N, SC.N)[..., 0]  # noqa: E731
    nr = np.linalg.norm(d["x_gt"])
    print("NRMSE vs ground truth: known-motion CG", np.linalg.norm(x_cg - d["x_gt"]) / nr,
          "pyramid (frame-aligned)", SC.aligned_nrmse(x_pyr, T_pyr, d["x_gt"], d["T_gt"], tr),
          "(opposite alignment)", SC.aligned_nrmse(x_pyr, -T_pyr + 2 * d["T_gt"], d["x_gt"], d["T_gt"], tr),
          "unaligned", np.linalg.norm(x_pyr - d["x_gt"]) / nr, flush=True)
    print(f"wrote {out} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
