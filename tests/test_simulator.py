"""Simulation-side operators (Simulation/encodeDISORDER.m, decodeDISORDER.m,
errorFit.m, precondDISORDER.m, Xsolver.m, gradientDISORDER.m,
directionDISORDER.m) and the joint pass of pyramidDISORDER.m:62-132.

CPU: the oracle's decode is the adjoint of its encode over n1 n2 (dot
test); errorFit vanishes on clean data and its per-shot values sum to the
total; Xsolver reaches the dense least-squares solution and counts its
operations like the MATLAB loop; the direction's first step has a largest
rotation of exactly 1 degree, is invariant to a common scale of gradient
and Hessian, and the weights follow the accept/reject rule; restrictTransform
zeroes what falls outside its limits; product direction/restrict = oracle.
The oracle's joint loop recovers the motion of a smooth volume.
GPU: encode, decode, errorFit, Xsolver, the gradient sums and three joint
passes against the oracle on a dense-mask 3-D problem, then convergence."""
import math

import numpy as np
import pytest

from disorder_amd import sampling as SA
from disorder_amd import simulator as SM
from oracle import simulator_oracle as SO
from tests.helpers import checkered_sampling, crandn, rel_l2, smooth_volume


def _problem(seed=0, n=(16, 12, 8), C=3, NS=4, tile=(2, 2)):
    rng = np.random.default_rng(seed)
    H = checkered_sampling(rng, n[0], n[1], NS, tile[0], tile[1])
    A = np.stack([(H == s + 1).astype(float) for s in range(NS)], axis=-1)
    S = crandn(rng, n + (C,))
    x = crandn(rng, n)
    T = np.zeros((NS, 6))
    T[:, :3] = rng.uniform(-1, 1, (NS, 3))
    T[:, 3:] = rng.uniform(-3, 3, (NS, 3)) * math.pi / 180
    return rng, H, A, S, x, T, n


def test_decode_is_scaled_adjoint_of_encode():
    rng, H, A, S, x, T, n = _problem()
    y = crandn(rng, n + (S.shape[3], A.shape[2])) * A[:, :, None, None, :]
    lhs = np.vdot(y, SO.encode_sim(x, S, A, T, n))
    rhs = n[0] * n[1] * np.vdot(SO.decode_sim(y, S, A, T, n), x)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_error_fit_clean_and_per_shot():
    rng, H, A, S, x, T, n = _problem(1)
    y = SO.encode_sim(x, S, A, T, n)
    assert SO.error_fit(x, y, S, A, T, n) < 1e-18 * np.sum(np.abs(y) ** 2)
    y2 = y + crandn(rng, y.shape) * A[:, :, None, None, :]
    per = SO.error_fit(x, y2, S, A, T, n, per_shot=True)
    assert abs(per.sum() - SO.error_fit(x, y2, S, A, T, n)) < 1e-9 * per.sum()


def test_xsolver_reaches_least_squares_and_counts():
    rng, H, A, S, x_gt, T, n = _problem(2, n=(6, 4, 2), C=2, NS=4)
    y = SO.encode_sim(x_gt, S, A, T, n) + 0.01 * crandn(rng, n + (2, 4)) * A[:, :, None, None, :]
    # dense operator: columns = encode of unit vectors (masked)
    N = int(np.prod(n))
    cols = []
    for i in range(N):
        e = np.zeros(N, complex)
        e[i] = 1
        cols.append(SO.encode_sim(e.reshape(n, order="F"), S, A, T, n).reshape(-1, order="F"))
    Mx = np.stack(cols, 1)
    ref = np.linalg.lstsq(Mx, (y * A[:, :, None, None, :]).reshape(-1, order="F"), rcond=None)[0]
    xs, nops = SO.xsolver(np.zeros(n, complex), y, S, A, T, n, nX=600, toler=1e-22)
    assert rel_l2(xs.reshape(-1, order="F"), ref) < 1e-6
    # operation count (Xsolver.m:41-90): 3 to start from a given x (1 from
    # zero), 2 per system application, 1 more per continued iteration; the
    # loop stops at the first check with n >= nX: 3, 5, 6, 8, 9, 11 -> 11
    for nX, want in ((10, 11), (11, 11), (12, 14), (30, 32)):
        assert SO.xsolver(np.zeros(n, complex), y, S, A, T, n, nX=nX, toler=0.0)[1] == want
    # from zero: 1, 3, 4, 6, 7, 9, 10, 12 -> checks at 3, 6, 9, 12
    for nX, want in ((10, 12), (11, 12), (25, 27)):
        assert SO.xsolver(None, y, S, A, T, n, nX=nX, toler=0.0)[1] == want


def test_direction_properties():
    rng = np.random.default_rng(5)
    NS = 7
    Jm = rng.normal(size=(NS, 9, 6))
    H = np.einsum("sia,sib->sab", Jm, Jm)
    dt = rng.normal(size=(6, NS))
    ddt = np.stack([H[:, a - 1, b - 1] for a, b in SO.HA_SIM])
    Tup, w, flagw, winic = SO.direction(dt, ddt)
    assert abs(np.abs(Tup[:, 3:6]).max() - math.pi / 180) < 1e-15
    assert np.all(w == 1e-3) and np.all(flagw == 0)
    Tup2, _, _, _ = SO.direction(3.0 * dt, 3.0 * ddt)
    assert np.allclose(Tup2, Tup, rtol=1e-9, atol=0)
    # second call: accepted shots divide w by 1.2, rejected multiply by 2
    fl = (np.arange(NS) % 2).astype(float)
    _, w2, _, _ = SO.direction(dt, ddt, w, fl, winic)
    assert np.allclose(w2, np.where(fl == 1, 1e-3 / 1.2, 2e-3))
    # the product implementation agrees to rounding
    for args in ((dt, ddt), (dt, ddt, w, fl, winic)):
        a = SO.direction(*args)
        b = SM.direction(*args)
        assert np.allclose(a[0], b[0], rtol=1e-12, atol=1e-18) and np.allclose(a[1], b[1]) and a[3] == b[3]


def test_restrict():
    T = np.array([[5.0, -2.0, 0.3, 3.2, -0.1, -3.5], [-7.9, 7.9, 0.0, 0.1, 0.2, 0.3]])
    lim = [[-8, 7], [-6, 5], [0, 0]]
    R = SO.restrict(T, lim)
    assert R.tolist() == [[5.0, -2.0, 0.0, 0.0, -0.1, 0.0], [-7.9, 0.0, 0.0, 0.1, 0.2, 0.3]]
    assert np.array_equal(SM.restrict(T, lim), R)


def _dev_problem(dev, seed=7):
    import torch
    import disorder_amd as D
    rng, H, A, S, x, T, n = _problem(seed, n=(32, 24, 16), C=3, NS=8, tile=(2, 4))
    k1, k2, off = SA.harmonic_sampling(H, 8)
    to = lambda a: D.colmajor(torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64))).to(dev))
    S32 = S.astype(np.complex64).astype(np.complex128)
    x32 = x.astype(np.complex64).astype(np.complex128)
    return rng, A, S32, x32, T, n, k1, k2, off, to


def _compact(dense, k1, k2, off):
    """Dense masked planes (n1,n2,n3,C,NS) -> compact rows (Ns, n3, C)."""
    rows = []
    for s in range(len(off) - 1):
        rows.append(dense[k1[off[s]:off[s + 1]], k2[off[s]:off[s + 1]], :, :, s])
    return np.concatenate(rows, 0)


@pytest.mark.gpu
def test_operators_match_oracle(dev):
    rng, A, S, x, T, n, k1, k2, off, to = _dev_problem(dev)
    op = SM.SimOperator(to(S), k1, k2, off)
    yd = SO.encode_sim(x, S, A, T, n)
    assert rel_l2(op.encode(to(x), T).cpu().numpy(), _compact(yd, k1, k2, off)) < 1e-5
    noisy = yd + 0.05 * crandn(rng, yd.shape) * A[:, :, None, None, :]
    noisy = noisy.astype(np.complex64).astype(np.complex128)
    yc = to(_compact(noisy, k1, k2, off))
    assert rel_l2(op.decode(yc, T).cpu().numpy(), SO.decode_sim(noisy, S, A, T, n)) < 1e-5
    e_g = op.error_fit(to(x), yc, T, per_state=True)
    e_r = SO.error_fit(x, noisy, S, A, T, n, per_shot=True)
    assert np.abs(e_g / e_r - 1).max() < 1e-4
    # gradient sums away from the optimum (at the true motion the gradient is
    # a cancelling sum of noise terms and float32 rounding dominates it)
    Tp = T.copy()
    Tp[:, :3] += rng.uniform(-0.3, 0.3, (T.shape[0], 3))
    Tp[:, 3:] += rng.uniform(-1, 1, (T.shape[0], 3)) * math.pi / 180
    EX, dt, ddt = op.partials(to(x), yc, Tp)
    EXr, dtr, ddtr = SO.gradient(x, noisy, S, A, Tp, n)
    assert np.abs(EX / EXr - 1).max() < 1e-4
    assert rel_l2(dt, dtr) < 1e-4, rel_l2(dt, dtr)
    assert rel_l2(ddt, ddtr) < 1e-4, rel_l2(ddt, ddtr)
    xg, ng = op.xsolver(yc, T, x=to(np.zeros(n)), nX=40)
    xr, nr = SO.xsolver(np.zeros(n, complex), noisy, S, A, T, n, nX=40)
    assert ng == nr and rel_l2(xg.cpu().numpy(), xr) < 1e-4, (ng, nr, rel_l2(xg.cpu().numpy(), xr))
    xg, ng = op.xsolver(yc, T, None, nX=25)
    xr, nr = SO.xsolver(None, noisy, S, A, T, n, nX=25)
    assert ng == nr and rel_l2(xg.cpu().numpy(), xr) < 1e-4
    op.close()


def _joint_problem(seed=8, n=(32, 24, 16), noise=0.01):
    """Smooth volume, 8 shots, +-1 voxel in-plane and +-1 degree motion; no
    through-plane translation (restrictTransform's [0 0] limit zeroes it)."""
    rng, H, A, S, _, T, n = _problem(seed, n=n, C=3, NS=8, tile=(2, 4))
    x = smooth_volume(rng, n, nblob=8)
    T[:, 3:] /= 3
    T[:, 2] = 0
    S = S.astype(np.complex64).astype(np.complex128)
    y = SO.encode_sim(x, S, A, T, n)
    sc = np.sqrt(np.sum(np.abs(y) ** 2) / (np.count_nonzero(A) * n[2] * S.shape[3]))
    y = y + noise * sc * crandn(rng, y.shape) * A[:, :, None, None, :]
    y = y.astype(np.complex64).astype(np.complex128)
    lim = [[-(n[0] // 2), n[0] - n[0] // 2 - 1], [-(n[1] // 2), n[1] - n[1] // 2 - 1], [0, 0]]
    return rng, H, A, S, y, T - T.mean(0), n, lim


def test_oracle_joint_loop_converges():
    """pyramidDISORDER's joint loop (one level, full resolution) recovers the
    motion up to its mean: 8 passes, 0.01 voxel and 0.05 degree."""
    _, _, A, S, y, Tgt, n, lim = _joint_problem()
    x, _ = SO.xsolver(np.zeros(n, complex), y, S, A, np.zeros_like(Tgt), n, nX=30)
    T, st, E = np.zeros_like(Tgt), {}, []
    for _ in range(8):
        x, T, _ = SO.joint_iteration(x, y, S, A, T, n, st, lim)
        E.append(SO.error_fit(x, y, S, A, T, n))
    assert np.abs(T - Tgt)[:, :3].max() < 0.01
    assert np.abs(T - Tgt)[:, 3:].max() * 180 / math.pi < 0.05
    assert E[-1] < 0.05 * E[0]


@pytest.mark.gpu
def test_joint_iterations_match_oracle(dev):
    """First three passes against the oracle step by step (per-shot
    acceptance, motion, image, operation count); then five more on the GPU
    alone reach the motion as the oracle does."""
    import torch
    import disorder_amd as D
    _, H, A, S, y, Tgt, n, lim = _joint_problem()
    k1, k2, off = SA.harmonic_sampling(H, 8)
    to = lambda a: D.colmajor(torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64))).to(dev))
    op = SM.SimOperator(to(S), k1, k2, off)
    yc = to(_compact(y, k1, k2, off))
    T0 = np.zeros_like(Tgt)
    xg, ng = op.xsolver(yc, T0, x=to(np.zeros(n)), nX=30)
    xr, nr = SO.xsolver(np.zeros(n, complex), y, S, A, T0, n, nX=30)
    assert ng == nr and rel_l2(xg.cpu().numpy(), xr) < 1e-4
    sg, sr = {}, {}
    Tg, Tr = T0.copy(), T0.copy()
    for it in range(3):
        xg, Tg, ng = SM.joint_iteration(op, xg, yc, Tg, sg, lim)
        xr, Tr, nr = SO.joint_iteration(xr, y, S, A, Tr, n, sr, lim)
        assert ng == nr, (it, ng, nr)
        assert np.array_equal(sg["flagw"], sr["flagw"]), (it, sg["flagw"], sr["flagw"])
        assert np.abs(Tg - Tr).max() < 1e-4, (it, np.abs(Tg - Tr).max())
        assert rel_l2(xg.cpu().numpy(), xr) < 1e-3, (it, rel_l2(xg.cpu().numpy(), xr))
    for _ in range(5):
        xg, Tg, _ = SM.joint_iteration(op, xg, yc, Tg, sg, lim)
    assert np.abs(Tg - Tgt)[:, :3].max() < 0.01
    assert np.abs(Tg - Tgt)[:, 3:].max() * 180 / math.pi < 0.05
    op.close()
