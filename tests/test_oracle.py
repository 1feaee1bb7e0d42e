"""Pins the FP64 oracle (CPU only).  The reference ships no tests, fixtures or
golden vectors, and MATLAB is not available, so every oracle function is
pinned by a closed form, a library routine or an invariant chosen so that a
plausible mistake (sign, index, transposed operand, dropped term) fails."""
import math

import numpy as np
import pytest

from oracle import disorder_oracle as O
from tests.helpers import checkered_sampling, crandn, random_motion, rel_l2, smooth_volume


def test_dftmtx_matches_numpy_fft():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 12, 16):
        x = crandn(rng, n)
        assert np.allclose(O.dftmtx(n) @ x, np.fft.fft(x), atol=1e-12)
        # inverse is conj(F)/n (build1DFTM.m:25,52)
        assert np.allclose(np.conj(O.dftmtx(n)) @ np.fft.fft(x) / n, x, atol=1e-12)


def test_wrap_helpers_closed_forms():
    a = np.array([0.0, np.pi, -np.pi, 3 * np.pi, -3 * np.pi, 4.0, -4.0, 2 * np.pi])
    w = O.wrap_to_pi(a)
    assert w[0] == 0 and w[1] == np.pi and w[2] == -np.pi
    assert np.isclose(w[3], np.pi) and np.isclose(w[4], -np.pi)  # -3pi -> -pi (negative odd multiple)
    assert np.isclose(w[5], 4.0 - 2 * np.pi) and np.isclose(w[6], -4.0 + 2 * np.pi)
    assert np.isclose(w[7], 0.0)
    h = O.wrap_to_pi_half(np.array([0.0, 1.0, 2.0, -2.0, np.pi / 2, 3.0]))
    assert np.allclose(h, [0.0, 1.0, 2.0 - np.pi, -2.0 + np.pi, np.pi / 2, 3.0 - np.pi])


def test_transform_grids_closed_form():
    r, k = O.transform_grids((8, 6, 5), (4, 6, 5))
    # r = (N/n)(j - ceil((n+1)/2)), j = 1..n
    assert np.allclose(r[0], 2.0 * (np.arange(1, 5) - 3))
    assert np.allclose(r[2], np.arange(1, 6) - 3)
    # k ifftshifted: FFT order with Nyquist at -n/2, scaled by 2pi/N (N full size)
    assert np.allclose(k[0], 2 * np.pi / 8 * np.array([0, 1, -2, -1]))
    assert np.allclose(k[2], 2 * np.pi / 5 * np.array([0, 1, 2, -2, -1]))


def test_identity_roll_unitary_inverse():
    rng = np.random.default_rng(1)
    n = (8, 6, 10)
    x = crandn(rng, n)
    T = np.array([[0, 0, 0, 0, 0, 0], [2, -1, 3, 0, 0, 0], [0.3, -0.7, 1.2, 0.2, -0.3, 0.1]])
    y = O.rigid_forward(x, T, n, n)
    assert np.abs(y[..., 0] - x).max() < 1e-13
    assert np.abs(y[..., 1] - np.roll(x, (2, -1, 3), axis=(0, 1, 2))).max() < 1e-12
    for s in range(3):
        assert abs(np.linalg.norm(y[..., s]) - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)
    back = O.rigid_inverse(y, T, n, n, sumR=False)
    for s in range(3):
        assert np.abs(back[..., s] - x).max() < 1e-12


def test_translation_sign_matches_shifting():
    """A fractional translation t along dim d equals shifting(x, {d: t})
    (Methods/shifting.m), an independent restatement of the Fourier shift."""
    rng = np.random.default_rng(2)
    n = (8, 6, 5)
    x = crandn(rng, n)
    for d in range(3):
        T = np.zeros((1, 6))
        T[0, d] = 0.37
        y = O.rigid_forward(x, T, n, n)[..., 0]
        z = O.shifting(x, {d: 0.37})
        assert np.abs(y - z).max() < 1e-12


def test_quarter_turn_is_permutation():
    """Rotation by pi/2 in a square plane maps grid points to grid points:
    tan(pi/4) = sin(pi/2) = 1 make every shear an integer circular shear,
    so the result is (to rounding) the torus rotation about the centre."""
    rng = np.random.default_rng(3)
    n = (8, 8, 4)
    x = crandn(rng, n)
    T = np.zeros((1, 6))
    T[0, 3] = np.pi / 2  # rotation in the (1,2) plane
    y = O.rigid_forward(x, T, n, n)[..., 0]
    # the sense derived in test_rotation_of_smooth_blob_matches_analytic:
    # y(r1, r2) = x(r2, -r1) about the centre; np.rot90(x, 1)[i, j] =
    # x[j, n - 1 - i] is that up to a circular shift; the other sense must
    # not fit
    for k, want in ((1, True), (-1, False)):
        cands = [np.roll(np.rot90(x, k, axes=(0, 1)), (s1, s2), axis=(0, 1)) for s1 in range(8) for s2 in range(8)]
        err = min(np.abs(y - c).max() for c in cands)
        assert (err < 1e-10) == want, (k, err)


def test_rotation_of_smooth_blob_matches_analytic():
    """Rotating a wide Gaussian blob by a small angle matches the analytically
    rotated blob (sinc interpolation is exact for band-limited content)."""
    n = (48, 48, 8)
    g = np.meshgrid(*[np.arange(1, v + 1) - math.ceil((v + 1) / 2) for v in n], indexing="ij")
    sig = 4.0
    c = (3.0, -2.0)

    def blob(a, b):
        return np.exp(-((a - c[0]) ** 2 + (b - c[1]) ** 2) / (2 * sig * sig)) * np.ones_like(g[2])

    x = blob(g[0], g[1]).astype(np.complex128)
    th = 0.2
    T = np.zeros((1, 6))
    T[0, 3] = th
    y = O.rigid_forward(x, T, n, n)[..., 0]
    err = {s: np.abs(y - blob(np.cos(s * th) * g[0] + np.sin(s * th) * g[1],
                              -np.sin(s * th) * g[0] + np.cos(s * th) * g[1])).max()
           for s in (1, -1)}
    # the sense follows from the shear factors of
    # precomputeFactorsSincRigidTransform.m:58-60 with di = 1 (sigma = -1):
    # e2 = exp(+i tan(th/2) r2 k1) moves dim 1 by +tan(th/2) r2
    # (y(r) = x(r1 + tan(th/2) r2, r2)), e3 = exp(-i sin(th) r1 k2) gives
    # y(r) = x(r1, r2 - sin(th) r1); shear-shear-shear composes to
    # y(r) = x(cos th r1 + sin th r2, -sin th r1 + cos th r2)
    assert err[1] < 1e-4 and err[-1] > 0.1, err


def _rot(g, a, b, th):
    """Coordinates after the rotation of plane (a, b) that the three shears
    of sincRigidTransform.m:42-54 apply: (r_a, r_b) -> (cos th r_a +
    sin th r_b, -sin th r_a + cos th r_b), as derived above for (1, 2)."""
    out = list(g)
    out[a] = np.cos(th) * g[a] + np.sin(th) * g[b]
    out[b] = -np.sin(th) * g[a] + np.cos(th) * g[b]
    return out


def test_rotation_composition_order_and_translation():
    """All three angles and a translation on an anisotropic blob against the
    analytic result.  sincRigidTransform.m:42-72 applies parameter 4 (plane
    (1,2): p1 = 1, p2 = 2), then 5 (plane (3,1)), then 6 (plane (2,3)),
    then the translation, so the output at r samples the input at
    R4 R5 R6 (r - t) (each R_m the coordinate map of _rot); the reversed
    order and the opposite sense both miss by far more than the tolerance."""
    n = (40, 40, 40)
    g = np.meshgrid(*[np.arange(1, v + 1) - math.ceil((v + 1) / 2) for v in n], indexing="ij")
    sig = (3.5, 4.5, 3.0)
    c = (2.0, -1.5, 1.0)

    def blob(r):
        return np.exp(-sum((r[d] - c[d]) ** 2 / (2 * sig[d] ** 2) for d in range(3)))

    x = blob(g).astype(np.complex128)
    t = np.array([0.7, -1.2, 0.4])
    th = (0.12, -0.09, 0.15)
    T = np.array([[*t, *th]])
    y = O.rigid_forward(x, T, n, n)[..., 0]
    r = [g[d] - t[d] for d in range(3)]

    def chain(order, sgn):
        q = r
        for m in order:  # innermost map first: R4 R5 R6 r applies R6 to r first
            a, b = ((0, 1), (2, 0), (1, 2))[m]
            q = _rot(q, a, b, sgn * th[m])
        return blob(q)

    good = np.abs(y - chain((2, 1, 0), 1)).max()
    assert good < 5e-4, good
    assert np.abs(y - chain((0, 1, 2), 1)).max() > 20 * good  # composition reversed
    assert np.abs(y - chain((2, 1, 0), -1)).max() > 0.05      # sense reversed


def test_flip_path_is_unitary_and_invertible():
    rng = np.random.default_rng(4)
    for n in ((8, 6, 10), (7, 9, 5)):
        x = crandn(rng, n)
        T = np.array([[0.1, 0.2, 0.3, 2.0, -2.5, 1.9]])
        y = O.rigid_forward(x, T, n, n)[..., 0]
        assert abs(np.linalg.norm(y) - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)
        back = O.rigid_inverse(y[..., None], T, n, n)
        assert np.abs(back - x).max() < 1e-12


def test_flipping_is_reflection_for_odd_sizes():
    """For odd n the flip shifts are integers (shifting.m:29-33): the result is
    an exact index reflection."""
    rng = np.random.default_rng(5)
    x = crandn(rng, (7, 5))
    y = O.flipping(x, {0: 7 / 2 + 1, 1: 5 / 2 + 1})
    # x3[j] = x[(n-1 - (j+h1 mod n) - h1) mod n], h1 = -(n+1)/2
    idx0 = [((7 - 1 - ((j - 4) % 7)) + 4) % 7 for j in range(7)]
    idx1 = [((5 - 1 - ((j - 3) % 5)) + 3) % 5 for j in range(5)]
    assert np.abs(y - x[np.ix_(idx0, idx1)]).max() < 1e-14


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    n = (8, 6, 10)
    x = smooth_volume(rng, n) + 0.05 * crandn(rng, n)
    T = np.array([[0.3, -0.7, 1.2, 0.2, -0.3, 0.1]])
    f = O.Factors(T, n, n, 1, cg=True)
    _, xB = O.sinc_rigid_transform(x, f, 1, keep=True)
    G = O.sinc_rigid_transform_gradient(xB, f)
    for g in range(6):
        h = 1e-6
        Tp, Tm = T.copy(), T.copy()
        Tp[0, g] += h
        Tm[0, g] -= h
        fd = (O.rigid_forward(x, Tp, n, n) - O.rigid_forward(x, Tm, n, n)) / (2 * h)
        assert rel_l2(G[g], fd) < 1e-6, g


def test_encode_rowdft_form_equals_fft_mask_form_and_adjoint():
    rng = np.random.default_rng(7)
    n, C, S = (8, 6, 5), 3, 4
    x = crandn(rng, n)
    Sm = crandn(rng, n + (C,))
    H = checkered_sampling(rng, 8, 6, 4, 2, 2)
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off[:S + 1], 8, 6, rng.uniform(0.5, 2, off[S]))
    T = random_motion(rng, S, 3, 1)
    y = O.encode(x, Sm, samp, T, n)
    assert np.abs(y - O.encode_rowdft(x, Sm, samp, T, n)).max() < 1e-12
    yr = crandn(rng, y.shape)
    samp_nw = O.Sampling(k1, k2, off[:S + 1], 8, 6)
    lhs = np.vdot(yr, y)
    rhs = np.vdot(O.decode(yr, Sm, samp_nw, T, n), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # weights are applied once in decode and in encodedecode (decode.m:46, encodedecode.m:133)
    ehe = O.encodedecode(x, Sm, samp, T, n)
    assert np.abs(ehe - O.decode(y, Sm, samp, T, n)).max() < 1e-12


def test_encode_unitary_dft_when_trivial():
    """T = 0, S = 1 coil of ones, full mask: E is the unitary 2D DFT over dims 1-2."""
    rng = np.random.default_rng(8)
    n = (4, 6, 3)
    x = crandn(rng, n)
    H = np.ones((4, 6), np.int64)
    k1, k2, off, _ = O.build_harmonic_sampling(H, 1)
    samp = O.Sampling(k1, k2, off[:2], 4, 6)
    y = O.encode(x, np.ones(n + (1,)), samp, np.zeros((1, 6)), n)
    X = np.fft.fft2(x, axes=(0, 1), norm="ortho")
    assert np.abs(y[:, :, 0] - X.reshape(24, 3, order="F")).max() < 1e-12


def test_build_harmonic_sampling_integer():
    H = np.array([[1, 2, 0], [2, 1, 1]])
    k1, k2, off, idx = O.build_harmonic_sampling(H, 2)
    # column-major linear order within each state, unsampled group last
    assert list(off) == [0, 3, 5, 6]
    assert list(zip(k1, k2)) == [(0, 0), (1, 1), (1, 2), (1, 0), (0, 1), (0, 2)]
    assert list(idx) == [0, 3, 5, 1, 2, 4]


def test_checkered_mask_cover():
    rng = np.random.default_rng(9)
    H = checkered_sampling(rng, 16, 12, 16, 4, 4)
    assert H.min() == 1 and H.max() == 16
    counts = np.bincount(H.ravel())[1:]
    assert np.all(counts == 16 * 12 // 16)
    # one sample per tile per state (tiles in the fftshifted frame)
    Hs = np.fft.fftshift(H)
    for o in range(4):
        for p in range(3):
            t = Hs[o * 4:(o + 1) * 4, p * 4:(p + 1) * 4]
            assert sorted(t.ravel()) == list(range(1, 17))


def test_cg_matches_dense_least_squares():
    rng = np.random.default_rng(10)
    n, C, S = (4, 4, 3), 2, 4
    Sm = crandn(rng, n + (C,))
    H = rng.integers(1, S + 1, size=(4, 4))
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off[:S + 1], 4, 4)
    T = random_motion(rng, S, 5, 1)
    x = crandn(rng, n) * 1e5  # CG stops on |rs| < 1e-6 in absolute terms (CGsolver.m:103)
    y = O.encode(x, Sm, samp, T, n)
    cols = []
    for i in range(x.size):
        e = np.zeros(x.size, complex)
        e[i] = 1
        cols.append(O.encode(e.reshape(n, order="C"), Sm, samp, T, n).ravel())
    A = np.stack(cols, 1)
    xl = np.linalg.lstsq(A, y.ravel(), rcond=None)[0].reshape(n)
    xc, En, nn = O.cg_solver(y, Sm, samp, T, n, nIt=100, tolType="RelativeResidual2Error", tol=1e-10)
    assert rel_l2(xc, xl) < 1e-8
    # early exit on zero data (CGsolver.m:70-75)
    x0, _, n0 = O.cg_solver(np.zeros_like(y), Sm, samp, T, n, nIt=5)
    assert n0 == 0 and np.all(x0 == 0)


def test_lm_partials_match_explicit_jacobian():
    rng = np.random.default_rng(11)
    n, C, S = (8, 6, 4), 2, 2
    x = smooth_volume(rng, n)
    Sm = crandn(rng, n + (C,))
    H = rng.integers(1, S + 1, size=(8, 6))
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off[:S + 1], 8, 6)
    T = random_motion(rng, S, 3, 1)
    y = O.encode(x, Sm, samp, random_motion(rng, S, 3, 1), n) + 0.01 * crandn(rng, (samp.Ns, n[2], C))
    En, dG, dH = O.lm_partials(y, x, Sm, samp, T, n)
    for s in range(S):
        sl = samp.state(s)[0]

        def res(Ts):
            TT = T.copy()
            TT[s] = Ts
            return (O.encode(x, Sm, samp, TT, n) - y)[sl].ravel()

        r0 = res(T[s])
        J = []
        for g in range(6):
            h = 1e-6
            e = np.zeros(6)
            e[g] = h
            J.append((res(T[s] + e) - res(T[s] - e)) / (2 * h))
        J = np.stack(J, 1)
        assert abs(En[s] - np.sum(np.abs(r0) ** 2) / y.size) < 1e-12
        assert np.allclose(dG[s], np.real(J.T @ np.conj(r0)), rtol=1e-6, atol=1e-9)
        JH = np.real(J.T @ np.conj(J))
        full = [JH[i, j] for (i, j) in O.HA]
        assert np.allclose(dH[s], full, rtol=1e-6, atol=1e-9)


def test_lm_recovers_known_motion():
    rng = np.random.default_rng(12)
    n, C, S = (12, 12, 6), 3, 4
    x = smooth_volume(rng, n, 8)
    Sm = 0.5 + 0.2 * crandn(rng, n + (C,))
    H = checkered_sampling(rng, 12, 12, 4, 2, 2)
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off[:S + 1], 12, 12)
    Tgt = random_motion(rng, S, 2, 0.5)
    y = O.encode(x, Sm, samp, Tgt, n)
    T = np.zeros((S, 6))
    w = np.full(S, 1e-3)
    cT = np.zeros(S, bool)
    for _ in range(25):
        T, w, cT, _ = O.lm_solver(y, x, Sm, samp, T, n, w, cT, 1e-3, (1e-6, 1e-6))
    assert np.abs(T[:, :3] - Tgt[:, :3]).max() < 1e-3
    assert np.abs(T[:, 3:] - Tgt[:, 3:]).max() * 180 / np.pi < 1e-3


def test_lm_step_closed_form():
    rng = np.random.default_rng(13)
    A = rng.standard_normal((6, 6))
    Hm = A @ A.T + np.eye(6)
    dH = np.array([Hm[i, j] for (i, j) in O.HA])
    g = rng.standard_normal(6)
    w = 0.3
    M = Hm.copy()
    M[np.diag_indices(6)] *= (1 + w)
    assert np.allclose(O.lm_step(dH, g, w, 1e-3), -(1e-3 / w) * np.linalg.solve(M, g))


def test_compress_coils_subspace():
    rng = np.random.default_rng(14)
    n, C = (6, 5, 4), 6
    base = crandn(rng, n + (3,))
    mix = crandn(rng, (3, C))
    Sm = base @ mix + 1e-3 * crandn(rng, n + (C,))
    Sc, _, P, A = O.compress_coils(Sm, 0.999)
    assert np.allclose(P, np.conj(P).T, atol=1e-10)
    ev = np.linalg.eigvalsh(0.5 * (P + np.conj(P).T))
    assert ev.min() > -1e-10
    assert A.shape[0] == 3
    # projector onto the kept subspace matches the top eigenvectors of P
    w, V = np.linalg.eigh(0.5 * (P + np.conj(P).T))
    Vk = V[:, ::-1][:, :3]
    assert np.allclose(np.conj(A).T @ A, Vk @ np.conj(Vk).T, atol=1e-8)


def test_resample_fourier_roundtrip():
    rng = np.random.default_rng(15)
    x = crandn(rng, (8, 6, 4))
    up = O.resample_fourier(x, (16, 12, 8))
    down = O.resample_fourier(up, (8, 6, 4))
    assert np.abs(down - x).max() < 1e-12
    assert O.level_size((256, 192, 256), 3) == [32, 24, 32]
    assert O.level_size((15, 9, 7), 1) == [8, 5, 4]


def test_outer_segments_use_untransformed_x():
    """Segments beyond NMs are encoded from the untransformed x (encode.m:76-79)."""
    rng = np.random.default_rng(16)
    n, C, S = (8, 6, 4), 2, 3
    x = crandn(rng, n)
    Sm = crandn(rng, n + (C,))
    H = rng.integers(1, S + 1, size=(8, 6))
    k1, k2, off, _ = O.build_harmonic_sampling(H, S)
    samp = O.Sampling(k1, k2, off[:S + 1], 8, 6)
    T = random_motion(rng, 2, 3, 1)
    y = O.encode(x, Sm, samp, T, n, nmotion=2)
    yfull = O.encode(x, Sm, samp, np.vstack([T, np.zeros((1, 6))]), n)
    assert np.abs(y - yfull).max() < 1e-12
    yr = crandn(rng, y.shape)
    assert abs(np.vdot(yr, y) - np.vdot(O.decode(yr, Sm, samp, T, n, nmotion=2), x)) < 1e-10 * abs(np.vdot(yr, y))


def test_lm_singular_system_takes_no_step_and_wrap_is_bounded():
    """A singular damped system gives no step (the reference's mldivide
    would return Inf/NaN, and its rotation wrap would then never end); far-out
    angles are wrapped in bounded time."""
    step = O.lm_step(np.zeros(21), np.ones(6), 1e-3, 1e-3)
    assert np.array_equal(step, np.zeros(6))
    T = np.array([[0, 0, 0, 1e30, -7.0, 4.0]])
    R = O.restrict_transform(T)
    assert np.all(np.abs(R[:, 3:]) <= np.pi)
    assert abs(R[0, 4] - (-7.0 + 2 * np.pi)) < 1e-15 and abs(R[0, 5] - (4.0 - 2 * np.pi)) < 1e-15


def test_dense_dft_mode_matches_fft_path():
    """set_dense_dft routes every DFT through a complex64 GEMM with
    dftmtx(n) (Methods/fftGPU.m): E^H E equals the FFT path to single
    precision, and a single axis DFT equals numpy's FFT."""
    from tests.helpers import crandn, rel_l2
    rng = np.random.default_rng(4)
    n, C = (16, 12, 10), 3
    x, Sm = crandn(rng, n), crandn(rng, n + (C,))
    H = rng.integers(1, 5, size=n[:2])
    k1, k2, off, _ = O.build_harmonic_sampling(H, 4)
    samp = O.Sampling(k1, k2, off[:5], n[0], n[1])
    T = np.array([[1.2, -0.7, 0.4, 0.03, -0.05, 0.02]] * 4)
    a = O.encodedecode(x, Sm, samp, T, n)
    O.set_dense_dft(True)
    try:
        b = O.encodedecode(x, Sm, samp, T, n)
        f = O.fft_dim(x, 1)
    finally:
        O.set_dense_dft(False)
    assert rel_l2(b, a) < 2e-6
    assert rel_l2(f, np.fft.fft(x, axis=1)) < 1e-6
