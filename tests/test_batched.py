"""Batched CG over independent reconstructions of uneven sizes with
hipGraph-captured iterations (dis_cg_batched).

GPU: the batched solver (graph and eager replay) against one dis_cg call per
volume with the same settings ('RelativeResidual2Error', tol 0, x0 = 0):
the same kernels run on the same data, so the results are bit-identical;
and one volume against the FP64 oracle's PCG (rel-L2 <= 1e-5).  The graph is
replayed in two separate calls with eager work in between, each checked."""
import numpy as np
import pytest

from disorder_amd import sampling as SA
from oracle import disorder_oracle as O
from tests.helpers import crandn, random_motion, rel_l2, smooth_volume

# phase-encode sizes on the register coil path (deterministic, no atomics):
# results are bit-identical however the iterations are launched
VOLS = [((32, 24, 10), (4, 3)), ((32, 20, 12), (4, 4)), ((64, 12, 8), (2, 3))]
# sizes served by the LDS coil kernel, whose adjoint adds with atomics: equal
# to FP32 summation-order differences
VOLS_GENERIC = [((16, 12, 10), (4, 3)), ((24, 18, 16), (2, 3))]


def test_truncate_shots_is_uneven_and_disjoint():
    from disorder_amd import synth
    rng = np.random.default_rng(0)
    perms = SA.permutation_table(rng, 4 * 4, 16)
    H = SA.random_checkered(16, 16, 4, 4, perms)
    Ht = synth.truncate_shots(H, 0.25, 3)
    counts = np.bincount(Ht.ravel(), minlength=17)[1:]
    assert (counts < 16).sum() == 4 and (counts == 16).sum() == 12
    assert np.all((Ht == H) | (Ht == 0))


@pytest.mark.gpu
def test_batched_matches_single_and_oracle(dev):
    import disorder_amd as D
    from disorder_amd import synth
    P = synth.batched_problem(dev, C=4, seed=3, volumes=VOLS, trunc=0.25)
    Es = [p["E"] for p in P]
    ys = [p["y"] for p in P]
    Ts = [p["T"] for p in P]
    Ses = [p["Se"] for p in P]
    nIt = 6
    xg, resg, msg = D.CGsolver_batched(ys, Es, Ts, Ses, nIt=nIt, use_graph=True)
    # eager work between the two graph runs
    singles = [D.CGsolver(y, E, T, x=None, nIt=nIt, tolType="RelativeResidual2Error", tol=0.0, Se=Se)[0]
               for y, E, T, Se in zip(ys, Es, Ts, Ses)]
    xg2, resg2, _ = D.CGsolver_batched(ys, Es, Ts, Ses, nIt=nIt, use_graph=True)
    xe, rese, _ = D.CGsolver_batched(ys, Es, Ts, Ses, nIt=nIt, use_graph=False)
    for i in range(len(P)):
        a, b, c, s = (t.cpu().numpy() for t in (xg[i], xg2[i], xe[i], singles[i]))
        assert np.array_equal(a, s), (i, np.abs(a - s).max())
        assert np.array_equal(b, s) and np.array_equal(c, s)
    assert np.array_equal(resg, rese) and np.array_equal(resg, resg2)
    assert np.all(np.diff(resg, axis=1) < 0)  # residual decreases every iteration
    # one volume against the FP64 oracle
    p = P[1]
    E = p["E"]
    samp = O.Sampling(E.k1, E.k2, E.off, p["n"][0], p["n"][1])
    Sm = p["Smaps"].cpu().numpy().astype(np.complex128)
    yn = p["y"].cpu().numpy().astype(np.complex128)
    Se = p["Se"].cpu().numpy().astype(np.float64)
    xo, _, _ = O.cg_solver(yn, Sm, samp, p["T"], p["n"], nIt=nIt, tolType="RelativeResidual2Error", tol=0.0, Se=Se)
    assert rel_l2(xg[1].cpu().numpy(), xo) < 1e-5


@pytest.mark.gpu
def test_batched_generic_sizes_match_single(dev):
    import disorder_amd as D
    from disorder_amd import synth
    P = synth.batched_problem(dev, C=4, seed=5, volumes=VOLS_GENERIC, trunc=0.25)
    Es, ys, Ts, Ses = [p["E"] for p in P], [p["y"] for p in P], [p["T"] for p in P], [p["Se"] for p in P]
    xg, resg, _ = D.CGsolver_batched(ys, Es, Ts, Ses, nIt=5, use_graph=True)
    for i, (y, E, T, Se) in enumerate(zip(ys, Es, Ts, Ses)):
        s = D.CGsolver(y, E, T, x=None, nIt=5, tolType="RelativeResidual2Error", tol=0.0, Se=Se)[0]
        assert rel_l2(xg[i].cpu().numpy(), s.cpu().numpy()) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(32, 24, 20), (20, 18, 14)])
def test_every_pass_shape_is_correct_and_autotune_picks_one(dev, n):
    """Each candidate shape the autotuner may pick (register path, LDS
    engine with radix 4/8/16 and 1..Pmax pencils) transforms correctly."""
    import ctypes
    import disorder_amd as D
    from tests.helpers import crandn, random_motion
    from tests.test_gpu_parity import to_dev, to_np
    rng = np.random.default_rng(sum(n))
    x = crandn(rng, n)
    T = random_motion(rng, 3)
    ref = O.rigid_forward(x, T, n, n)
    ctx = D.default_context(0)
    nn = (ctypes.c_int32 * 3)(*n)
    shapes = [(1, P, r) for r in (4, 8, 16) for P in (1, 2, 8, 0)]
    for g, P, r in shapes:
        for axis in range(3):
            D._lib.check(ctx.lib.dis_set_pass_shape(ctx.h, nn, axis, g, P, r), ctx.h)
        out = D.sincRigidTransform(to_dev(x, dev), T, di=1)
        assert rel_l2(to_np(out), ref) < 1e-5, (g, P, r)
    D._lib.check(ctx.lib.dis_autotune_clear(ctx.h), ctx.h)
    choice = D.autotune(n, S=3)
    assert len(choice) == 3 and all(c["ms"] > 0 for c in choice)
    out = D.sincRigidTransform(to_dev(x, dev), T, di=1)
    assert rel_l2(to_np(out), ref) < 1e-5
    D._lib.check(ctx.lib.dis_autotune_clear(ctx.h), ctx.h)


@pytest.mark.gpu
@pytest.mark.parametrize("n,tiles,C", [((32, 24, 10), (4, 3), 4), ((64, 40, 6), (2, 2), 8), ((160, 16, 4), (4, 2), 3)])
def test_persistent_work_queue_matches_static_grid(dev, n, tiles, C):
    """The persistent (shot, readout-plane) work-queue launch of the coil
    stage gives bit-identical encode / decode / E^H E / energies to the
    one-workgroup-per-item launch on uneven (truncated) shots."""
    import ctypes
    import disorder_amd as D
    from disorder_amd import synth
    P = synth.batched_problem(dev, C=C, seed=7, volumes=[(n, tiles)], trunc=0.5)[0]
    E, y, T, x = P["E"], P["y"], P["T"], P["x_gt"]
    assert P["lines"].max() > 1.25 * P["lines"].min()
    ctx = E.ctx
    out = {}
    for mode in (1, 2):
        D._lib.check(ctx.lib.dis_set_coil_schedule(ctx.h, mode), ctx.h)
        out[mode] = (E.encode(x, T).cpu().numpy(), E.decode(y, T).cpu().numpy(),
                     E.encodedecode(x, T).cpu().numpy(), E.energy(x, y, T, per_sample=True))
    D._lib.check(ctx.lib.dis_set_coil_schedule(ctx.h, 0), ctx.h)
    for a, b in zip(out[1], out[2]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("budget_vols", [1, 2, 3])
def test_state_batches_match_one_launch_per_pass(dev, budget_vols):
    """Transforms run state batch by state batch (the cache budget holding
    1, 2 or 3 volumes of a 48x40x24 grid, 7 states, so batches of uneven
    size) give encode, decode, E^H E, energies and the rigid transform
    bit-identical to every state in every pass launch (budget 0); flips
    (rotations beyond 90 degrees) run unbatched."""
    import disorder_amd as D
    from tests.test_gpu_parity import to_dev
    n, C, S = (48, 40, 24), 3, 7
    rng = np.random.default_rng(17 + budget_vols)
    x = smooth_volume(rng, n) + 0.1 * crandn(rng, n)
    Sm = crandn(rng, n + (C,)) * 0.5
    H = rng.integers(1, S + 1, size=(n[0], n[1]))
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    T = random_motion(rng, S)
    T[2, 3] = 2.2  # a rotation beyond 90 degrees: that state's transform flips
    ctx = D.Context(0)
    E = D.Encoding(to_dev(Sm, dev), k1, k2, off[:S + 1], ctx=ctx)
    xd = to_dev(x, dev)
    y = E.encode(xd, T)
    out = {}
    for budget in (0, budget_vols * 8 * n[0] * n[1] * n[2]):
        ctx.set_cache_budget(budget)
        for Tc in (T, np.where(np.arange(S)[:, None] == 2, 0.0, T)):  # with and without the flip
            key = (budget, Tc is T)
            out[key] = (E.encode(xd, Tc).cpu().numpy(), E.decode(y, Tc).cpu().numpy(),
                        E.encodedecode(xd, Tc).cpu().numpy(), E.energy(xd, y, Tc, per_sample=True),
                        D.sincRigidTransform(xd, Tc, di=1, ctx=ctx).cpu().numpy())
    ctx.set_cache_budget(208 << 20)
    for flip in (True, False):
        for a, b in zip(out[(0, flip)], out[(budget_vols * 8 * n[0] * n[1] * n[2], flip)]):
            assert np.array_equal(a, b), flip
    E.close()


@pytest.mark.gpu
def test_grouped_profile_counts_every_launch(dev):
    """Profiling E^H E with one state per batch (cache budget of one volume)
    counts S times the pass launches of the unbatched run (budget 0), the
    same algorithmic bytes and the same single broadcast and accumulate
    launch per transform, although the batched passes of each transform share
    one event pair; every class's time is positive."""
    import disorder_amd as D
    from tests.test_gpu_parity import to_dev
    n, C, S = (40, 32, 16), 2, 5
    rng = np.random.default_rng(23)
    x = smooth_volume(rng, n)
    Sm = crandn(rng, n + (C,)) * 0.5
    H = rng.integers(1, S + 1, size=(n[0], n[1]))
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    T = random_motion(rng, S)
    ctx = D.Context(0)
    E = D.Encoding(to_dev(Sm, dev), k1, k2, off[:S + 1], ctx=ctx)
    xd = to_dev(x, dev)
    prof = {}
    for budget in (0, 8 * n[0] * n[1] * n[2]):
        ctx.set_cache_budget(budget)
        E.encodedecode(xd, T)  # warm-up outside the profile
        ctx.profile(True)
        E.encodedecode(xd, T)
        prof[budget] = ctx.profile_read()
        ctx.profile(False)
    ctx.set_cache_budget(208 << 20)
    flat, one = prof[0], prof[8 * n[0] * n[1] * n[2]]
    assert flat["pass"][0] > 0 and one["pass"][0] == S * flat["pass"][0]
    assert one["pass"][2] == pytest.approx(flat["pass"][2], rel=1e-12)
    for k in ("bcast_pass", "accum_pass", "coil_stage"):
        assert one[k][0] == flat[k][0] >= 1, k
        assert one[k][2] == pytest.approx(flat[k][2], rel=1e-12), k
    for p in (flat, one):
        assert all(v[1] > 0 for v in p.values() if v[0]), p
    E.close()
