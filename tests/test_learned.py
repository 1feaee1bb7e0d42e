"""Learned motion proposal: the network's hand-written HIP forward, backward
and Adam kernels against the FP64 CPU oracle (oracle/nn_oracle.py), whose
backward is itself pinned by central finite differences.

Tolerances: the device rounds both operands of every matrix product to
bfloat16 and accumulates in FP32; the oracle applies the same rounding and
sums exactly.  Most outputs then agree to FP32 accumulation error (median
relative error <= 1e-6), but an intermediate activation computed in FP32 on
the device and in FP64 in the oracle can sit on either side of a bf16
rounding boundary (probability ~1e-7 / 2^-8 per value); such a flip moves one
operand by one bf16 ulp, so the whole-array bounds are rel-L2 <= 2e-4
(outputs) and <= 5e-4 (gradients)."""
import math

import numpy as np
import pytest

from oracle import nn_oracle as NO
from tests.helpers import rel_l2


def _params(seed, scale=1.0):
    from disorder_amd.learned import MotionNet
    p = MotionNet.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    return p + scale * 0.05 * rng.standard_normal(p.size)  # non-zero biases, perturbed conv


def _batch(seed, lens=(7, 3, 12)):
    rng = np.random.default_rng(seed)
    so = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    X = rng.standard_normal((so[-1], NO.NIN))
    return X, so


def test_oracle_backward_matches_finite_differences():
    p = _params(3)
    X, so = _batch(4, (4, 2))
    rng = np.random.default_rng(5)
    dY = rng.standard_normal((so[-1], NO.NOUT))
    g, dX = NO.backward(p, X, so, dY, bf16=False)

    def f(pp, XX):
        return float(np.sum(NO.forward(pp, XX, so, bf16=False)[0] * dY))
    h = 1e-6
    idx = np.concatenate([rng.choice(NO.NPARAMS, 40, replace=False),
                          np.arange(NO.NPARAMS - 186, NO.NPARAMS)])  # random entries + all conv parameters
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (f(p + e, X) - f(p - e, X)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    for r, c in [(0, 0), (3, 12), (5, 7)]:
        E = np.zeros_like(X)
        E[r, c] = h
        fd = (f(p, X + E) - f(p, X - E)) / (2 * h)
        assert abs(fd - dX[r, c]) <= 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -8 + 2 ** -20], np.float32)
    r = NO.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie -> even
    assert r[2] == 1.0 + 4 * 2 ** -8                # tie -> even (up)
    assert r[3] == -2.5
    assert r[4] == 1.0 + 2 ** -7                    # above the tie -> up


def test_features_diagonal_newton_step():
    from disorder_amd.learned import features
    S = 5
    rng = np.random.default_rng(2)
    h = rng.uniform(1, 10, (S, 6))
    d = rng.normal(0, 1, (S, 6))          # true update (vox, rad)
    d[:, 3:] *= 0.1                       # within the +-20 degree clip
    g = -h * d                            # gradient of a diagonal quadratic at the current point
    sums = np.zeros((S, 28))
    sums[:, 0] = rng.uniform(1, 2, S)
    sums[:, 1:7] = g
    sums[:, 7:13] = h
    F = features(sums)
    assert np.allclose(F[:, :3], d[:, :3], rtol=1e-5)
    assert np.allclose(F[:, 3:6], d[:, 3:] * 180 / math.pi, rtol=1e-5)
    assert np.allclose(F[:, 6:12].mean(axis=0), np.log(h / h.mean(0)).mean(0), atol=1e-5)
    # a near-flat direction (tiny curvature) gives a step clipped to +-NEWTON_CLIP
    sums[0, 7] = 1e-12
    sums[0, 1] = -1.0
    assert features(sums)[0, 0] == pytest.approx(20.0)


def test_relative_target_and_proposal_centring():
    """The training target is the true update less its mean over shots (a
    common offset of the estimate is not in it), and the network learns the
    target less the Newton step of the features."""
    from disorder_amd.learned import net_target, relative_target
    rng = np.random.default_rng(4)
    S = 9
    T_true = rng.normal(0, 1, (S, 6))
    e = rng.normal(0, 0.3, (S, 6))
    t1 = relative_target(T_true, T_true + e)
    t2 = relative_target(T_true, T_true + e + rng.normal(0, 5, (1, 6)))  # plus a common offset
    np.testing.assert_allclose(t1, t2, atol=1e-4)
    np.testing.assert_allclose(t1.mean(axis=0), 0.0, atol=1e-5)
    np.testing.assert_allclose(t1[:, :3], -(e[:, :3] - e[:, :3].mean(0)), atol=1e-5)
    np.testing.assert_allclose(t1[:, 3:], -(e[:, 3:] - e[:, 3:].mean(0)) * 180 / math.pi, atol=1e-4)
    F = rng.normal(0, 1, (S, 13)).astype(np.float32)
    np.testing.assert_allclose(net_target(F, t1) + F[:, :6], t1, atol=1e-6)


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("lens", [(7, 3, 12), (1,), (64, 48, 33, 5)])
def test_forward_backward_match_oracle(dev, lens):
    import torch
    from disorder_amd.learned import MotionNet
    p = _params(11)
    X, so = _batch(12, lens)
    rng = np.random.default_rng(13)
    dY = rng.standard_normal((so[-1], NO.NOUT))
    net = MotionNet(p.astype(np.float32))
    Xd = torch.from_numpy(X.astype(np.float32)).to(dev)
    Y = net.forward(Xd, so).cpu().numpy()
    p32 = p.astype(np.float32).astype(np.float64)
    X32 = X.astype(np.float32).astype(np.float64)
    Yr, _ = NO.forward(p32, X32, so, bf16=True)
    assert rel_l2(Y, Yr) < 2e-4
    assert np.median(np.abs(Y - Yr) / (np.abs(Yr) + 1e-3)) < 1e-6
    G, dX = net.backward(Xd, so, torch.from_numpy(dY.astype(np.float32)).to(dev))
    Gr, dXr = NO.backward(p32, X32, so, dY.astype(np.float32).astype(np.float64), bf16=True)
    G = G.cpu().numpy()
    q, qr = NO.unpack(G), NO.unpack(Gr)
    for name, _ in NO.SHAPES:
        assert rel_l2(q[name], qr[name]) < 5e-4, name
    assert rel_l2(dX.cpu().numpy(), dXr) < 5e-4
    # the bf16 products differ from the plain FP64 network by ~2^-9, far above the bounds above
    Y64, _ = NO.forward(p32, X32, so, bf16=False)
    assert rel_l2(Y, Y64) > 5 * rel_l2(Y, Yr)


@pytest.mark.gpu
def test_adam_step_matches_oracle(dev):
    import torch
    from disorder_amd.learned import MotionNet
    p = _params(21)
    X, so = _batch(22, (10, 6))
    rng = np.random.default_rng(23)
    tgt = rng.standard_normal((so[-1], NO.NOUT))
    net = MotionNet(p.astype(np.float32))
    Xd = torch.from_numpy(X.astype(np.float32)).to(dev)
    Td = torch.from_numpy(tgt.astype(np.float32)).to(dev)
    loss = net.train_step(Xd, Td, so, lr=1e-3)
    p32 = p.astype(np.float32).astype(np.float64)
    X32 = X.astype(np.float32).astype(np.float64)
    T32 = tgt.astype(np.float32).astype(np.float64)
    Y, _ = NO.forward(p32, X32, so, bf16=True)
    n = Y.size
    assert abs(loss - np.mean((Y - T32) ** 2)) < 1e-5 * max(1.0, loss)
    g, _ = NO.backward(p32, X32, so, 2 * (Y - T32) / n, bf16=True)
    pn, _, _ = NO.adam_step(p32, g, np.zeros_like(g), np.zeros_like(g), 1, 1e-3)
    # the first Adam step moves every parameter by ~lr * sign(g); compare the moves
    mv, mvr = net.p.cpu().numpy() - p32, pn - p32
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert np.allclose(mv[big], mvr[big], atol=2e-6)


@pytest.mark.gpu
def test_training_reduces_loss(dev):
    """Adam on a smooth synthetic regression (target = 0.5 * first six
    features plus the mean of the neighbours) lowers the loss by 5x in 40
    epochs of 64 scans."""
    import torch
    from disorder_amd.learned import MotionNet, train
    rng = np.random.default_rng(31)
    data = []
    for _ in range(64):
        S = int(rng.integers(16, 40))
        F = rng.standard_normal((S, NO.NIN)).astype(np.float32)
        t = 0.5 * F[:, :6]
        t[1:-1] += 0.25 * (F[:-2, :6] + F[2:, :6])
        data.append((F, t.astype(np.float32)))
    net = MotionNet(seed=0)
    losses = train(net, data, epochs=40, batch=512, lr=1e-3, residual=False)
    assert losses[-1] < 0.2 * losses[0], losses[::8]
    assert all(b <= a * 1.05 for a, b in zip(losses[::8], losses[8::8]))


@pytest.mark.gpu
def test_trained_proposal_against_plain_pyramid(dev, capsys):
    """The committed trained network inside the joint pyramid on the
    reference's 64^3 case (the inputs of tests/golden/small_case.py: 8
    coils, 16 shots, motion U(+-5 deg, +-3 voxel), 30 dB, 3 levels), plain
    and with the proposer (first two outer iterations of every level, no
    retirement).  Measured and printed, by the learned-proposal criterion:
    outer iterations, LM trials and solve time until the motion error (common
    motion removed) first lies below 0.1 voxel and 0.1 degree, and the final
    frame-aligned NRMSE against the ground truth.  Asserted: proposals are
    kept for some shots (the comparison is not vacuous), both runs reach the
    motion criterion, and learned+LM ends within 1e-3 of the plain NRMSE
    (a kept proposal moves the path of every later CG and LM step; the
    energy test keeps it from degrading the fit)."""
    import os
    import time
    import torch
    import disorder_amd as D
    from disorder_amd import learned as LN
    from disorder_amd import pyramid as PY
    from tests.golden import small_case as SC
    assert os.path.exists(LN.WEIGHTS), "trained weights missing"
    d = SC.make_inputs()

    def to_dev(a):
        return D.colmajor(torch.from_numpy(np.asfortranarray(a.astype(np.complex64))).to(dev))
    y, Sm = to_dev(d["y"]), to_dev(d["Sm"])

    def tr(v, Tm):
        return D.sincRigidTransform(to_dev(v), Tm, di=1)[..., 0].cpu().numpy()

    def run(prop):
        hit = {}
        count = {"outer": 0}

        def cb(li, it, x, T, t):
            count["outer"] += 1
            e = LN.relative_error(T, d["T_gt"])
            if "t" not in hit and np.abs(e[:, :3]).max() < 0.1 and np.abs(e[:, 3:]).max() < 0.1:
                hit.update(t=t, outer=count["outer"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, T, info = PY.solveXT(y, Sm, d["k1"], d["k2"], d["off"], SC.N, callback=cb, proposer=prop, **SC.PYR)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return {"nrmse": SC.aligned_nrmse(x.cpu().numpy(), T, d["x_gt"], d["T_gt"], tr), "wall": wall,
                "outer": [lv["outer"] for lv in info["levels"]], "lm_trials": info["lm_trials"],
                "t_crit": hit.get("t"), "outer_crit": hit.get("outer")}
    def _s(t):
        return "never" if t is None else f"{t:.3f} s"
    plain = run(None)
    prop = LN.Proposer(LN.MotionNet.load(), retire=False)
    learned = run(prop)
    with capsys.disabled():
        print(f"\n[learned proposal, 64^3 reference case] kept {prop.accepted}/{prop.proposed}; "
              f"plain: outer {plain['outer']}, LM trials {plain['lm_trials']}, motion < 0.1 after "
              f"{plain['outer_crit']} outer / {_s(plain['t_crit'])}, NRMSE {plain['nrmse']:.5f}, "
              f"{plain['wall']:.3f} s; learned: outer {learned['outer']}, LM trials {learned['lm_trials']}, motion "
              f"< 0.1 after {learned['outer_crit']} outer / {_s(learned['t_crit'])}, NRMSE {learned['nrmse']:.5f}, "
              f"{learned['wall']:.3f} s")
    assert prop.accepted > 0
    assert plain["t_crit"] is not None and learned["t_crit"] is not None
    assert learned["nrmse"] <= plain["nrmse"] + 1e-3


def test_proposer_retires_after_a_round_that_keeps_nothing(monkeypatch):
    """Host logic of Proposer(retire=True) with the GPU calls stubbed: a
    round in which some shot's energy drops keeps proposing; the first round
    in which no shot keeps its proposal retires the proposer, and later calls
    (other levels too) return the motion unchanged without evaluating
    anything."""
    import types
    import torch
    from disorder_amd import learned as LN
    S = 4
    calls = {"sums": 0, "energy": 0}
    drop = {"on": True}

    def fake_sums(y, E, x, T):
        calls["sums"] += 1
        s = np.zeros((S, 28))
        s[:, 0] = 1.0
        s[:, 7:13] = 1.0
        return s

    def fake_energy(E, x, y, T, off, Mbe=None):
        calls["energy"] += 1
        return np.full(S, 0.5 if drop["on"] else 2.0)

    class FakeNet:
        def forward(self, F, off):
            return torch.zeros((F.shape[0], 6))

    monkeypatch.setattr(LN, "lm_sums", fake_sums)
    monkeypatch.setattr(LN, "state_energy", fake_energy)
    monkeypatch.setattr(LN, "level_scale", lambda L: 1.0)
    P = LN.Proposer(FakeNet(), max_outer=2)
    L1 = types.SimpleNamespace(y=None, E=None, off=np.arange(S + 1))
    L2 = types.SimpleNamespace(y=None, E=None, off=np.arange(S + 1))
    T = np.zeros((S, 6))
    cT = np.zeros(S, bool)
    _, take = P(L1, torch.zeros(1), T, cT)
    assert take.all() and not P.retired and P.accepted == S
    drop["on"] = False
    _, take = P(L1, torch.zeros(1), T, cT)
    assert not take.any() and P.retired
    n = dict(calls)
    Tn, take = P(L2, torch.zeros(1), T, cT)  # a new level: still retired
    assert not take.any() and np.array_equal(Tn, T) and calls == n and P.proposed == 2 * S
    # without retirement the new level proposes again
    Q = LN.Proposer(FakeNet(), max_outer=2, retire=False)
    Q(L1, torch.zeros(1), T, cT)
    Q(L2, torch.zeros(1), T, cT)
    assert Q.proposed == 2 * S and not Q.retired
    # a new solve (pyramid.solveXT calls reset first) proposes again
    P.reset()
    drop["on"] = True
    _, take = P(L1, torch.zeros(1), T, cT)
    assert take.all() and not P.retired and P.proposed == 3 * S
