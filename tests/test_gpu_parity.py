"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle on the
same seeded inputs.  Tolerance: relative L2 <= 1e-5 for the transform and the
encode/decode operators (FP32 complex arithmetic vs FP64)."""
import numpy as np
import pytest
import torch

from oracle import disorder_oracle as O
from tests.helpers import checkered_sampling, crandn, random_motion, rel_l2, smooth_volume

pytestmark = pytest.mark.gpu
TOL = 1e-5


def to_dev(a, dev):
    import disorder_amd as D
    t = torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64)))
    return D.colmajor(t.to(dev))


def to_np(t):
    return t.cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("n,Nfull", [((16, 12, 10), None), ((16, 12, 10), (32, 24, 20)), ((15, 14, 9), None),
                                     ((28, 20, 21), (56, 40, 42)), ((64, 48, 64), None),
                                     # register-path lengths 32, 64, 128, 192, 256 on every axis
                                     ((32, 64, 128), (64, 128, 256)), ((192, 16, 32), None), ((256, 32, 20), None),
                                     ((24, 256, 16), None), ((16, 18, 192), None),
                                     # radix-10 and radix-14 register paths (160, 224)
                                     ((160, 20, 16), None), ((20, 224, 12), (40, 224, 24)), ((16, 18, 160), None),
                                     ((224, 16, 10), None),
                                     # odd radix-9 register path (144) on every axis
                                     ((144, 12, 10), None), ((16, 144, 12), (32, 144, 24)), ((12, 16, 144), None),
                                     # radix-6 and odd radix-3 register paths (96, 48) on every axis
                                     ((96, 16, 12), None), ((16, 96, 48), (32, 192, 96)), ((48, 12, 96), None),
                                     # prime factors 11 and 13 (runtime-radix stage): 11, 13, 22, 26, 176
                                     # on every axis, and a prime length 97
                                     ((11, 13, 22), None), ((26, 11, 13), (52, 22, 26)), ((22, 26, 11), None),
                                     ((176, 12, 10), None), ((12, 176, 11), (24, 176, 22)), ((13, 10, 176), None),
                                     ((97, 6, 8), None)])
def test_rigid_forward_inverse(dev, n, Nfull):
    import disorder_amd as D
    rng = np.random.default_rng(sum(n))
    Nf = Nfull or n
    x = crandn(rng, n)
    S = 4
    T = random_motion(rng, S)
    T[1] = 0.0
    xt = D.sincRigidTransform(to_dev(x, dev), T, di=1, Nfull=Nf)
    ref = O.rigid_forward(x, T, n, Nf)
    for s in range(S):
        assert rel_l2(to_np(xt)[..., s], ref[..., s]) < TOL, (n, s)
    # inverse with the state sum and without
    v = crandn(rng, n + (S,))
    out = D.sincRigidTransform(to_dev(v, dev), T, di=0, sumR=True, Nfull=Nf)
    assert rel_l2(to_np(out), O.rigid_inverse(v, T, n, Nf, sumR=True)) < TOL
    outs = D.sincRigidTransform(to_dev(v, dev), T, di=0, sumR=False, Nfull=Nf)
    refs = O.rigid_inverse(v, T, n, Nf, sumR=False)
    for s in range(S):
        assert rel_l2(to_np(outs)[..., s], refs[..., s]) < TOL


@pytest.mark.parametrize("n", [(16, 12, 10), (15, 9, 14)])
def test_rigid_flip_path(dev, n):
    """Rotations beyond 90 degrees (Methods/flipping.m via sincRigidTransform.m:43,109)."""
    import disorder_amd as D
    rng = np.random.default_rng(7)
    x = crandn(rng, n)
    T = np.array([[0.3, -0.2, 0.5, 2.2, 0.1, -0.05],
                  [0.0, 0.4, 0.0, 0.1, -2.5, 0.2],
                  [0.2, 0.0, -0.3, -0.1, 0.05, 1.9],
                  [0.1, 0.1, 0.1, 0.05, 0.02, 0.03]])
    xt = to_np(D.sincRigidTransform(to_dev(x, dev), T, di=1))
    ref = O.rigid_forward(x, T, n, n)
    for s in range(T.shape[0]):
        assert rel_l2(xt[..., s], ref[..., s]) < TOL, s
    v = crandn(rng, n + (T.shape[0],))
    assert rel_l2(to_np(D.sincRigidTransform(to_dev(v, dev), T, di=0)), O.rigid_inverse(v, T, n, n)) < TOL


def test_rigid_properties_full_size(dev):
    """Size-independent properties at the config grid 256x192x256: T=0 is the
    identity, an integer translation is a circular shift, inverse(forward) = I,
    unitarity."""
    import disorder_amd as D
    n = (256, 192, 256)
    g = torch.Generator(device=dev).manual_seed(3)
    x = D.colmajor(torch.randn(n, dtype=torch.complex64, device=dev, generator=g))
    T = np.array([[0, 0, 0, 0, 0, 0], [3, -2, 5, 0, 0, 0], [1.3, -0.7, 2.1, 0.05, -0.08, 0.06]], np.float64)
    xt = D.sincRigidTransform(x, T, di=1)
    nx = torch.linalg.vector_norm(x).item()
    e0 = (torch.linalg.vector_norm(xt[..., 0] - x) / nx).item()
    assert e0 < 1e-6
    rolled = torch.roll(x, shifts=(3, -2, 5), dims=(0, 1, 2))
    assert (torch.linalg.vector_norm(xt[..., 1] - rolled) / nx).item() < 1e-5
    for s in range(3):
        assert abs(torch.linalg.vector_norm(xt[..., s]).item() / nx - 1) < 1e-5
    back = D.sincRigidTransform(D.colmajor(xt[..., 2:3].contiguous()), T[2:3], di=0)
    assert (torch.linalg.vector_norm(back - x) / nx).item() < 1e-5


def test_rigid_rotation_sense_and_order_against_analytic(dev):
    """The HIP transform on an anisotropic Gaussian blob with all three
    angles and a translation equals the analytically moved blob: output at
    r = input at R4 R5 R6 (r - t), R_m rotating plane (1,2), (3,1), (2,3)
    as (r_a, r_b) -> (cos th r_a + sin th r_b, -sin th r_a + cos th r_b)
    (sincRigidTransform.m:42-72 with the factors of
    precomputeFactorsSincRigidTransform.m:58-60); the reversed order and the
    opposite sense miss."""
    import math
    import disorder_amd as D
    n = (64, 48, 40)
    g = np.meshgrid(*[np.arange(1, v + 1) - math.ceil((v + 1) / 2) for v in n], indexing="ij")
    sig, c = (4.0, 4.5, 3.5), (2.0, -1.5, 1.0)

    def blob(r):
        return np.exp(-sum((r[d] - c[d]) ** 2 / (2 * sig[d] ** 2) for d in range(3)))

    t, th = np.array([0.7, -1.2, 0.4]), (0.12, -0.09, 0.15)
    y = to_np(D.sincRigidTransform(to_dev(blob(g), dev), np.array([[*t, *th]]), di=1))[..., 0]

    def chain(order, sgn):
        q = [g[d] - t[d] for d in range(3)]
        for m in order:
            a, b = ((0, 1), (2, 0), (1, 2))[m]
            qa, qb = q[a], q[b]
            q = list(q)
            q[a] = np.cos(sgn * th[m]) * qa + np.sin(sgn * th[m]) * qb
            q[b] = -np.sin(sgn * th[m]) * qa + np.cos(sgn * th[m]) * qb
        return blob(q)

    good = np.abs(y - chain((2, 1, 0), 1)).max()
    assert good < 5e-4, good
    assert np.abs(y - chain((0, 1, 2), 1)).max() > 10 * good
    assert np.abs(y - chain((2, 1, 0), -1)).max() > 0.05


def _problem(rng, n, C, S, nmotion=None, weights=False, Nfull=None):
    x = smooth_volume(rng, n) + 0.1 * crandn(rng, n)
    Sm = crandn(rng, n + (C,)) * 0.5
    if S == 4:
        H = checkered_sampling(rng, n[0], n[1], 4, 2, 2)
    else:
        H = rng.integers(1, S + 1, size=(n[0], n[1]))
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    We = rng.uniform(0.5, 1.5, off[S]) if weights else None
    samp = O.Sampling(k1, k2, off[:S + 1], n[0], n[1], We)
    nm = S if nmotion is None else nmotion
    T = random_motion(rng, nm)
    return x, Sm, samp, T, nm, (Nfull or n)


@pytest.mark.parametrize("n,C,S,nmotion,weights", [((16, 12, 8), 3, 4, None, False),
                                                    ((16, 12, 8), 3, 5, 4, True),
                                                    ((20, 15, 12), 9, 6, None, True),
                                                    # register-path coil stage (n1 = 32, 64, 192, 256)
                                                    ((32, 24, 16), 3, 4, None, True),
                                                    ((64, 20, 8), 8, 5, 4, True),
                                                    ((192, 16, 6), 2, 4, None, False),
                                                    ((144, 18, 6), 3, 4, None, True),
                                                    ((256, 12, 4), 10, 4, None, True),
                                                    # states of ~900 and ~1280 samples: several sample
                                                    # chunks per workgroup on the register path
                                                    ((64, 28, 4), 8, 2, None, True),
                                                    ((64, 40, 4), 8, 2, None, False),
                                                    # one state of 3,072 samples: four adjoint table blocks
                                                    ((64, 48, 4), 3, 1, None, True),
                                                    # register coil stage at n1 = 160 and 224
                                                    ((160, 16, 4), 4, 4, None, True),
                                                    ((224, 12, 4), 3, 3, None, False),
                                                    # register coil stage at n1 = 96 and 48
                                                    ((96, 20, 8), 3, 4, None, True),
                                                    ((48, 24, 6), 2, 5, None, False),
                                                    # prime factors 11 and 13 on every axis (runtime-radix
                                                    # stage in the passes and the generic coil stage)
                                                    ((11, 13, 22), 3, 5, None, True),
                                                    ((26, 22, 13), 2, 4, 3, False),
                                                    ((176, 12, 11), 2, 4, None, True),
                                                    ((13, 176, 6), 2, 5, None, False),
                                                    ((12, 10, 176), 2, 4, None, True)])
def test_encode_decode_ehe(dev, n, C, S, nmotion, weights):
    import disorder_amd as D
    rng = np.random.default_rng(11)
    x, Sm, samp, T, nm, Nf = _problem(rng, n, C, S, nmotion, weights)
    E = D.Encoding(to_dev(Sm, dev), samp.k1, samp.k2, samp.off, Nfull=Nf, We=samp.We, nmotion=nm)
    Tall = np.vstack([T, np.zeros((S - nm, 6))])
    yr = O.encode(x, Sm, samp, T, Nf, nmotion=nm)
    xd = to_dev(x, dev)
    y = to_np(E.encode(xd, T))
    assert rel_l2(y, yr) < TOL
    yin = crandn(rng, yr.shape)
    xr = O.decode(yin, Sm, samp, T, Nf, nmotion=nm)
    assert rel_l2(to_np(E.decode(to_dev(yin, dev), T)), xr) < TOL
    er = O.encodedecode(x, Sm, samp, T, Nf, nmotion=nm)
    assert rel_l2(to_np(E.encodedecode(xd, T)), er) < TOL
    # residual energy per sample and total (FP64 accumulation)
    ps = E.energy(xd, to_dev(yin, dev), T, per_sample=True)
    pr = O.compute_energy(yin, x, Sm, samp, T, Nf, nmotion=nm, per_sample=True)
    assert rel_l2(ps, pr) < TOL
    del Tall


def test_adjoint_lane_table_edge_states(dev):
    """The adjoint's lane table (samples sorted by phase-encode row, dealt in
    equal runs to the lanes, rows spanning lanes finished from carries) on
    states that stress it: no samples, one sample, one row holding all 48 of
    its samples (a row spread over 48 lanes), 768 samples (three per lane),
    ten full rows plus 33 others (heavy rows, partly empty last lanes) and
    257 samples (two per lane); decode and E^H E against the oracle."""
    import disorder_amd as D
    rng = np.random.default_rng(29)
    n, C = (64, 48, 4), 2
    grid = np.array([(a, b) for a in range(n[0]) for b in range(n[1])])
    states = [np.zeros((0, 2), int), grid[rng.choice(len(grid), 1, replace=False)],
              np.array([(7, b) for b in range(n[1])]),
              grid[rng.choice(len(grid), 768, replace=False)]]
    full = np.array([(a, b) for a in range(10) for b in range(n[1])])
    rest = grid[grid[:, 0] >= 10]
    states.append(np.vstack([full, rest[rng.choice(len(rest), 33, replace=False)]]))
    states.append(grid[rng.choice(len(grid), 257, replace=False)])
    for st in states:
        rng.shuffle(st)  # acquisition order within a state is arbitrary
    k1 = np.concatenate([st[:, 0] for st in states]).astype(np.int64)
    k2 = np.concatenate([st[:, 1] for st in states]).astype(np.int64)
    off = np.concatenate([[0], np.cumsum([len(st) for st in states])])
    S = len(states)
    We = rng.uniform(0.5, 1.5, off[-1])
    samp = O.Sampling(k1, k2, off, n[0], n[1], We)
    x = smooth_volume(rng, n) + 0.1 * crandn(rng, n)
    Sm = crandn(rng, n + (C,)) * 0.5
    T = random_motion(rng, S)
    E = D.Encoding(to_dev(Sm, dev), k1, k2, off, We=We)
    yin = crandn(rng, (off[-1], n[2], C))
    xr = O.decode(yin, Sm, samp, T, n)
    assert rel_l2(to_np(E.decode(to_dev(yin, dev), T)), xr) < TOL
    er = O.encodedecode(x, Sm, samp, T, n)
    assert rel_l2(to_np(E.encodedecode(to_dev(x, dev), T)), er) < TOL


def test_encode_zero_motion_and_level_grid(dev):
    """T == 0 skips the transform (encode.m:61); level grids (Nfull = 2n)."""
    import disorder_amd as D
    rng = np.random.default_rng(5)
    n = (16, 12, 10)
    x, Sm, samp, T, nm, _ = _problem(rng, n, 2, 4)
    Nf = (32, 24, 20)
    E = D.Encoding(to_dev(Sm, dev), samp.k1, samp.k2, samp.off, Nfull=Nf)
    xd = to_dev(x, dev)
    for TT in (np.zeros((4, 6)), T):
        assert rel_l2(to_np(E.encode(xd, TT)), O.encode(x, Sm, samp, TT, Nf)) < TOL
        assert rel_l2(to_np(E.encodedecode(xd, TT)), O.encodedecode(x, Sm, samp, TT, Nf)) < TOL


def test_ehe_properties_config_grid(dev):
    """At 256x192x256, 8 coils, 64 DISORDER shots: the adjoint identity
    <E x, y> = <x, E^H y>, E^H E = decode(encode(x)), Hermitian E^H E."""
    import disorder_amd as D
    n = (256, 192, 256)
    C, S = 8, 64
    rng = np.random.default_rng(0)
    H = checkered_sampling(rng, n[0], n[1], S, 8, 8)
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    off = off[:S + 1]
    g = torch.Generator(device=dev).manual_seed(1)
    Sm = D.colmajor(torch.randn(n + (C,), dtype=torch.complex64, device=dev, generator=g) * 0.3)
    E = D.Encoding(Sm, k1, k2, off)
    T = random_motion(rng, S, rot_deg=3, tra=2)
    x = D.colmajor(torch.randn(n, dtype=torch.complex64, device=dev, generator=g))
    z = D.colmajor(torch.randn(n, dtype=torch.complex64, device=dev, generator=g))
    y = E.encode(x, T)
    yr = D.colmajor(torch.randn(tuple(y.shape), dtype=torch.complex64, device=dev, generator=g))
    a = torch.vdot(yr.flatten(), y.flatten()).item()
    b = torch.vdot(E.decode(yr, T).flatten(), x.flatten()).item()
    assert abs(a - b) / abs(a) < 1e-5
    ehe = E.encodedecode(x, T)
    dec = E.decode(y, T)
    assert (torch.linalg.vector_norm(ehe - dec) / torch.linalg.vector_norm(dec)).item() < 1e-5
    p = torch.vdot(z.flatten(), ehe.flatten()).item()
    q = torch.vdot(E.encodedecode(z, T).flatten(), x.flatten()).item()
    assert abs(p - q) / abs(p) < 1e-5


def test_ehe_two_shots_config_grid_matches_oracle(dev):
    """E^H E of two shots of the 64-shot DISORDER pattern (768 lines each) at
    the full 256x192x256 grid, 8 coils, against the FP64 oracle (rel-L2 <=
    1e-5): an error common to E and E^H passes the property checks above but
    not this one.  Smooth maps and a smooth volume plus noise, motion of a
    few degrees and voxels (one shot with a rotation beyond 90 degrees would
    not be a config case; both stay below)."""
    import disorder_amd as D
    n = (256, 192, 256)
    C, S = 8, 64
    rng = np.random.default_rng(3)
    H = checkered_sampling(rng, n[0], n[1], S, 8, 8)
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    pick = [5, 41]
    sl = [np.arange(off[s], off[s + 1]) for s in pick]
    k1s, k2s = np.concatenate([k1[i] for i in sl]), np.concatenate([k2[i] for i in sl])
    offs = np.array([0, len(sl[0]), len(sl[0]) + len(sl[1])])
    T = np.array([[1.3, -2.2, 0.7, 0.04, -0.03, 0.05], [-0.6, 1.1, 2.4, -0.05, 0.06, -0.02]])
    g = torch.Generator(device=dev).manual_seed(11)
    from disorder_amd import synth
    Sm = synth.coil_maps(n, C, seed=12, device=dev)
    x = D.colmajor(synth.phantom(n, 12, seed=13, device=dev)
                   + 0.05 * torch.randn(n, dtype=torch.complex64, device=dev, generator=g))
    E = D.Encoding(Sm, k1s, k2s, offs)
    out = E.encodedecode(x, T).cpu().numpy()
    yg = E.encode(x, T).cpu().numpy()
    E.close()
    xh = x.cpu().numpy().astype(np.complex128)
    Sh = Sm.cpu().numpy().astype(np.complex128)
    del Sm, x
    torch.cuda.empty_cache()
    O.set_fft_workers(16)
    try:
        samp = O.Sampling(k1s, k2s, offs, n[0], n[1])
        yr = O.encode(xh, Sh, samp, T, n)
        assert rel_l2(yg, yr) < TOL
        ref = O.decode(yr, Sh, samp, T, n)
    finally:
        O.set_fft_workers(None)
    assert rel_l2(out, ref) < TOL, rel_l2(out, ref)


@pytest.mark.parametrize("tolType,nIt,lam,use_pre", [("RelativeResidual2Error", 6, 0.0, True),
                                                     ("Energy", 7, 0.05, True), ("None", 4, 0.0, False)])
def test_cg_matches_oracle(dev, tolType, nIt, lam, use_pre):
    """CGsolver: iterates, energies and iteration count against the FP64 oracle
    (FP32 vectors, FP64 scalars: tolerance 1e-4 after a few iterations)."""
    import disorder_amd as D
    rng = np.random.default_rng(21)
    n, C, S = (16, 12, 8), 3, 4
    x, Sm, samp, T, nm, Nf = _problem(rng, n, C, S, weights=True)
    y = O.encode(x, Sm, samp, T, Nf) + 0.01 * crandn(rng, (samp.Ns, n[2], C))
    Se = 1.0 / (np.sum(np.abs(Sm) ** 2, axis=3) + 1e-9) if use_pre else None
    Ma = (rng.uniform(size=n) > 0.1).astype(np.float64)
    xr, Er, nr = O.cg_solver(y, Sm, samp, T, Nf, nIt=nIt, tolType=tolType, tol=0.0, Se=Se, Ma=Ma,
                             lam=(lam or None))
    E = D.Encoding(to_dev(Sm, dev), samp.k1, samp.k2, samp.off, Nfull=Nf, We=samp.We)

    def fdev(a):
        return None if a is None else D.colmajor(torch.from_numpy(np.asfortranarray(a.astype(np.float32))).to(dev))

    xg, Eg, ng = D.CGsolver(to_dev(y, dev), E, T, nIt=nIt, tolType=tolType, tol=0.0, Se=fdev(Se), Ma=fdev(Ma),
                            lam=lam)
    assert ng == nr
    assert rel_l2(to_np(xg), xr) < 1e-4
    if tolType == "Energy":
        assert Eg.shape == Er.shape
        assert np.allclose(Eg, Er, rtol=1e-4)
    # warm start from the oracle's solution continues identically
    xg2, _, _ = D.CGsolver(to_dev(y, dev), E, T, x=to_dev(xr, dev), nIt=2, tolType="None", Se=fdev(Se), Ma=fdev(Ma),
                           lam=lam)
    xr2, _, _ = O.cg_solver(y, Sm, samp, T, Nf, x0=xr, nIt=2, tolType="None", Se=Se, Ma=Ma, lam=(lam or None))
    assert rel_l2(to_np(xg2), xr2) < 1e-4


@pytest.mark.parametrize("n,Nfull,flip", [((16, 12, 10), None, False), ((32, 64, 16), (64, 128, 32), False),
                                          ((15, 12, 14), None, True), ((32, 16, 12), None, True),
                                          ((144, 12, 10), None, False)])
def test_gradient_matches_oracle(dev, n, Nfull, flip):
    """sincRigidTransformGradient: every d(T x)/dT_g against the FP64 oracle."""
    import disorder_amd as D
    rng = np.random.default_rng(31)
    Nf = Nfull or n
    x = smooth_volume(rng, n) + 0.05 * crandn(rng, n)
    T = random_motion(rng, 3, 4, 1.5)
    if flip:
        T[0, 3] = 2.1
        T[2, 5] = -2.4
    Tx, G = D.sincRigidTransformGradient(to_dev(x, dev), T, Nfull=Nf)
    Tx, G = to_np(Tx), to_np(G)
    for s in range(3):
        f = O.Factors(T[s:s + 1], n, Nf, 1, cg=True)
        xt, xB = O.sinc_rigid_transform(x, f, 1, keep=True)
        Gr = O.sinc_rigid_transform_gradient(xB, f)
        assert rel_l2(Tx[..., s], xt[..., 0]) < TOL
        for g in range(6):
            assert rel_l2(G[..., s, g], Gr[g][..., 0]) < 2e-5, (s, g)


@pytest.mark.parametrize("n,C,S", [((16, 12, 8), 3, 4), ((32, 24, 10), 4, 4)])
@pytest.mark.parametrize("fat", [None, "H", "F"])
def test_lm_partials_and_step_match_oracle(dev, n, C, S, fat):
    """LMsolver: per-state energy, gradient and Hessian sums, then one full
    accept/reject update, against the FP64 oracle; with a readout ROI E.nF
    (LMsolver.m:33,68,98) from fat-shift direction 'H' or 'F' the sums run
    over that readout slab only (the (32, 24, 10) grid on the register coil
    path, whose launches then cover only the slab's planes)."""
    import disorder_amd as D
    rng = np.random.default_rng(41)
    x, Sm, samp, _, _, Nf = _problem(rng, n, C, S)
    Tgt = random_motion(rng, S, 3, 1)
    y = O.encode(x, Sm, samp, Tgt, Nf) + 0.01 * crandn(rng, (samp.Ns, n[2], C))
    T0 = Tgt + random_motion(rng, S, 1.0, 0.3)
    E = D.Encoding(to_dev(Sm, dev), samp.k1, samp.k2, samp.off, Nfull=Nf)
    nF = None if fat is None else O.readout_roi(n[2], fat)
    E.set_readout_roi(nF)
    xd, yd = to_dev(x, dev), to_dev(y, dev)
    En, dG, dH = D.lm_partials(yd, E, xd, T0)
    Er, dGr, dHr = O.lm_partials(y, x, Sm, samp, T0, Nf, nF=nF)
    assert rel_l2(En, Er) < 1e-5
    assert rel_l2(dG, dGr) < 1e-4
    assert rel_l2(dH, dHr) < 1e-4
    if nF is not None:  # the slab changes the sums: the ROI is not a no-op
        Ef, _, _ = O.lm_partials(y, x, Sm, samp, T0, Nf)
        assert rel_l2(En, Ef) > 0.05
    w = np.full(S, 1e-3)
    cT = np.zeros(S, bool)
    cT[1] = True
    Tg, wg, cg, Eg, _ = D.LMsolver(yd, E, xd, T0, w, cT, winit=1e-3, tL=(0.05, 0.02))
    Tr, wr, cr, Ern = O.lm_solver(y, x, Sm, samp, T0, Nf, w, cT, 1e-3, (0.05, 0.02), nF=nF)
    assert np.allclose(wg, wr) and np.array_equal(cg, cr)
    assert np.abs(Tg - Tr).max() < 1e-5
    assert np.allclose(Eg, Ern, rtol=1e-5, atol=1e-12)
    assert np.all(Tg[1] == T0[1])  # converged states are untouched
