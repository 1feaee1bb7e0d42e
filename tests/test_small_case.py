"""The reference's CPU-runnable case on the GPU against the FP64 oracle's run
of the same case (tests/golden/small_case_64.npz, written by
tests/golden/make_small_case.py from the inputs of tests/golden/small_case.py):
64^3, 8 coils, 16 DISORDER shots of 4 x 4 random-checkered tiles, motion
U(+-5 deg, +-3 vox), 30 dB; 20 CG iterations with the known motion, then the
3-level joint pyramid (nX CG schedule, readout ROI 'H', closing 'Energy' CG).

Tolerances (the reconstruction bounds of the method): image rel-L2 to the
oracle's image <= 1e-4 for both steps, motion within 0.01 voxel and 0.01 deg,
the same number of outer iterations per level.  Against the ground truth
(printed): the joint estimate's motion error once the common motion is
removed, and its frame-aligned NRMSE (the image moved by the mean motion
offset, as Simulation/pyramidDISORDER.m:124-131 removes it) beside the
known-motion CG's.  The joint run's data pass the Gibbs-ringing Tukey window
of every level (downsampleOperators.m:54-57), which softens the phantom's
hard edges, so its NRMSE is bounded by twice the known-motion CG's; the
motion error by 0.1 voxel and 0.2 degree."""
import math
import os

import numpy as np
import pytest

from tests.golden import small_case as SC
from tests.helpers import rel_l2

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_case_64.npz")


def test_fixture_inputs_are_consistent():
    """CPU: the fixture belongs to these inputs (seed, shapes) and the
    known-motion CG image is closer to the ground truth than the zero image
    by a wide margin (a fixture of other inputs fails this)."""
    ref = np.load(FIXTURE, allow_pickle=False)
    assert int(ref["seed"]) == SC.SEED
    assert ref["x_cg"].shape == SC.N and ref["x_pyr"].shape == SC.N and ref["T_pyr"].shape == (SC.S, 6)
    x_gt = SC.phantom(SC.N, 10, SC.SEED + 1)
    assert np.linalg.norm(ref["x_cg"] - x_gt) / np.linalg.norm(x_gt) < 0.1


@pytest.mark.gpu
def test_small_case_matches_oracle_run(dev, capsys):
    import torch
    import disorder_amd as D
    from disorder_amd import pyramid as PY
    ref = np.load(FIXTURE, allow_pickle=False)
    d = SC.make_inputs()

    def to_dev(a):
        return D.colmajor(torch.from_numpy(np.asfortranarray(a.astype(np.complex64))).to(dev))
    y, Sm = to_dev(d["y"]), to_dev(d["Sm"])
    E = D.Encoding(Sm, d["k1"], d["k2"], d["off"])
    xk, _, nn = D.CGsolver(y, E, d["T_gt"], nIt=SC.CG_KNOWN, tolType="RelativeResidual2Error", tol=0.0)
    E.close()
    xk = xk.cpu().numpy()
    assert nn == SC.CG_KNOWN
    e_cg = rel_l2(xk, ref["x_cg"])
    assert e_cg < 1e-4, e_cg
    outer = {}

    def cb(li, it, x, T, t):
        outer[li] = max(outer.get(li, 0), it)
    x, T, info = PY.solveXT(y, Sm, d["k1"], d["k2"], d["off"], SC.N, callback=cb, **SC.PYR)
    x = x.cpu().numpy()
    e_pyr = rel_l2(x, ref["x_pyr"])
    dv = np.abs(T[:, :3] - ref["T_pyr"][:, :3]).max()
    dd = np.abs(T[:, 3:] - ref["T_pyr"][:, 3:]).max() * 180 / math.pi
    assert [lv["outer"] for lv in info["levels"]] == ref["outer"].tolist()
    assert e_pyr < 1e-4, e_pyr
    assert dv < 0.01 and dd < 0.01, (dv, dd)

    def tr(v, Tm):
        return D.sincRigidTransform(to_dev(v), Tm, di=1)[..., 0].cpu().numpy()
    nr = np.linalg.norm(d["x_gt"])
    cg_gt = np.linalg.norm(xk - d["x_gt"]) / nr
    pyr_gt = SC.aligned_nrmse(x, T, d["x_gt"], d["T_gt"], tr)
    Te = T - d["T_gt"]
    Te -= Te.mean(axis=0, keepdims=True)
    with capsys.disabled():
        print(f"\n[64^3 case] GPU vs oracle: CG rel-L2 {e_cg:.2e}, pyramid rel-L2 {e_pyr:.2e}, motion "
              f"{dv:.1e} vox / {dd:.1e} deg; NRMSE vs ground truth: known-motion CG {cg_gt:.4f}, joint "
              f"(frame-aligned) {pyr_gt:.4f}; motion error {np.abs(Te[:, :3]).max():.3f} vox / "
              f"{np.abs(Te[:, 3:]).max() * 180 / math.pi:.3f} deg")
    assert pyr_gt < 2.0 * cg_gt, (pyr_gt, cg_gt)
    assert np.abs(Te[:, :3]).max() < 0.1 and np.abs(Te[:, 3:]).max() * 180 / math.pi < 0.2
