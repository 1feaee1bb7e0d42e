"""Literal FP64 restatement of the compressed-sensing final reconstruction
(Reconstruction/CSsolver.m, the shearlet term of Reconstruction/regularize.m,
Reconstruction/traceEstimation.m) and of the shearlet system it builds
(Shearlets/buildShearlet.m and helpers) for the configuration CSsolver uses:
buildShearlet(N, J=1, gpu, K=1, 'kos') with the default 'maxflat2' filter,
full = 0, real shearlets.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
disorder_amd.shearlet and the GPU regulariser; the product never imports it.

Functions follow the MATLAB text step by step (1-based index arithmetic kept
where it matters); arrays are complex128 / float64.
Pinning (tests/test_shearlet.py): every helper (upsampling, centred
crop/pad, shearing with integer shifts = per-row circular shifts) against
its definition on small arrays; closed properties of the planes (low-pass DC
gain 1 and Hermitian, band-pass DC gain 0 so every directional band and
every 3-D band but the low-pass removes DC, cone 2 of an (a, b) plane = the
transposed cone 1 of the (b, a) plane with the shearing negated); the
regulariser is self-adjoint and positive, and with unit weights equals
IFFT3(sum_w |S_w|^2 FFT3 x); with unit weights the trace estimates converge
to their closed forms sum_w sum_k |S_w(k)|^2 and sum_c sum_v |S_c(v)|^2.
The plane filters themselves are parity unpinned against the reference:
without MATLAB no reference output exists, and the checks above do not fix
every coefficient.
"""
from __future__ import annotations

import math

import numpy as np

# ----------------------------------------------------------------------------
# Filters (Shearlets/buildDirectionalFilter.m 'kos', buildQuadratureMirrorFilter.m 'maxflat2')
# ----------------------------------------------------------------------------


def directional_filter_kos():
    a0 = -math.sqrt(3.0)
    a1 = -math.sqrt(3.0)
    a2 = 2.0 + math.sqrt(3.0)
    h0 = np.array([[0.0, -a1, -a0 * a1, 0.0],
                   [-a2, -a0 * a2, -a0, 1.0],
                   [0.0, a0 * a1 * a2, -a1 * a2, 0.0]])
    return h0 * (math.sqrt(2.0) / h0.sum())


MAXFLAT2 = [.0104933261758410, -.0263483047033631, -.0517766952966370, .276348304703363, .582566738241592,
            .276348304703363, -.0517766952966369, -.0263483047033631, .0104933261758408]


def qmf_maxflat2():
    f = np.array(MAXFLAT2)
    return f / f.sum()


def mirror_modulation(x):
    """Methods/mirrorModulation.m for a row vector: times (-1)^(j - floor(L/2) - 1)."""
    x = np.asarray(x, np.float64)
    L = x.shape[-1]
    c = L // 2 + 1
    return x * np.array([(-1.0) ** (j - c) for j in range(1, L + 1)])


# ----------------------------------------------------------------------------
# Helpers (Methods/upsampling.m, resampling.m fo=2, shearing.m)
# ----------------------------------------------------------------------------


def upsampling(x, dim, ups=2):
    """upsampling(x, dim, ups) with odd = 1: size ups*N - (ups-1), x at 1:ups:end."""
    x = np.asarray(x)
    N = list(x.shape)
    N[dim] = ups * N[dim] - (ups - 1)
    out = np.zeros(N, dtype=x.dtype)
    idx = [slice(None)] * x.ndim
    idx[dim] = slice(0, N[dim], ups)
    out[tuple(idx)] = x
    return out


def resample_fo2(x, Nres):
    """resampling(x, Nres, 2): centred zero-pad or crop per dimension in the
    array's own domain (no transforms): kept span orig..fina (1-based) about
    zeroF = ceil((Nmax+1)/2), orig = zeroF - ceil((Nmin-1)/2),
    fina = zeroF + floor((Nmin-1)/2)."""
    x = np.asarray(x)
    for m in range(len(Nres)):
        Nor, Nr = x.shape[m], Nres[m]
        if Nor == Nr:
            continue
        Nmin, Nmax = min(Nor, Nr), max(Nor, Nr)
        zeroF = math.ceil((Nmax + 1) / 2)
        orig = zeroF - math.ceil((Nmin - 1) / 2)
        fina = zeroF + (Nmin - 1) // 2
        if Nor < Nr:
            shp = list(x.shape)
            shp[m] = Nr
            out = np.zeros(shp, dtype=x.dtype)
            idx = [slice(None)] * x.ndim
            idx[m] = slice(orig - 1, fina)
            out[tuple(idx)] = x
            x = out
        else:
            idx = [slice(None)] * x.ndim
            idx[m] = slice(orig - 1, fina)
            x = x[tuple(idx)]
    return x


def fft2D(x):
    for m in range(2):
        x = np.fft.fft(np.fft.ifftshift(x, axes=m), axis=m)
    return x


def ifft2D(x):
    for m in range(2):
        x = np.fft.fftshift(np.fft.ifft(x, axis=m), axes=m)
    return x


def shearing(x, k):
    """shearing(x, k, [2 1]): row i (1-based) shifted along dim 2 by
    -k (i - floor(N1/2) - 1), by the Fourier phase of shifting.m:39-45."""
    N1, N2 = x.shape
    r = np.array([i - (N1 // 2 + 1) for i in range(1, N1 + 1)], np.float64)
    H = -k * r  # shift per row
    kg = (2 * np.pi / N2) * (np.arange(1, N2 + 1) - math.ceil((N2 + 1) / 2))
    kg = -1j * np.fft.ifftshift(kg)
    ph = np.exp(H[:, None] * kg[None, :])
    return np.fft.ifft(np.fft.fft(x, axis=1) * ph, axis=1)


# ----------------------------------------------------------------------------
# Shearlet system (Shearlets/*.m)
# ----------------------------------------------------------------------------


def feasible(Na, Nb):
    """checkFilterFeasibility.m configuration 1 ('kos', 'maxflat2') for K = 1."""
    L = 9
    M = directional_filter_kos().shape
    lcheck1 = L
    lcheck2 = L + 2 * (L - 1) + (M[0] - 1) * 2 ** 2
    return lcheck1 <= min(Na, Nb) and lcheck2 <= min(Na, Nb) and M[1] <= min(Na, Nb)


def wedge_bandpass_lowpass(N):
    """buildWedgeBandpassAndLowpassFilters.m for J = 1, K = [1], complex = 0.
    Returns (WD [N1, N2, 5] for k = 1, BP [N1, N2], LP [N1, N2])."""
    DF = directional_filter_kos()
    SF1 = qmf_maxflat2()
    WF = mirror_modulation(SF1)
    SF2 = SF1.copy()
    DF = DF / np.sum(np.abs(DF))
    FHP = WF
    FLP1 = SF1
    FLP2 = {2: SF2}
    FLP2[1] = np.convolve(FLP2[2], upsampling(FLP2[2][None, :], 1)[0])
    BP = fft2D(resample_fo2(FHP[None, :], N))
    k = 1
    WD = np.zeros((N[0], N[1], 2 ** (k + 1) + 1), np.complex128)
    DFups = upsampling(DF, 0, 2 ** (k + 1))
    from scipy.signal import convolve2d
    c2 = convolve2d(DFups, FLP2[2 - k][:, None])
    WDups = upsampling(resample_fo2(c2, N), 1, 2 ** k)
    FLPaux = resample_fo2(FLP2[2 - max(k - 1, 0)][None, :], WDups.shape)
    WDups = ifft2D(fft2D(WDups) * fft2D(FLPaux))
    FLPauxfl = FLPaux[:, ::-1]
    for l in range(-2 ** k, 2 ** k + 1):
        W = shearing(WDups, l)
        W = ifft2D(fft2D(W) * fft2D(FLPauxfl))
        W = (2 ** k) * W[:, 0:(2 ** k) * N[1]:2 ** k]
        WD[:, :, 2 ** k + 1 - l - 1] = fft2D(W)
    LP = fft2D(resample_fo2(np.outer(FLP1, FLP1), N))
    return WD, BP, LP


def shearlets_2d(Na, Nb):
    """getShearlets(prepareFilters([Na Nb]), [Na Nb], idxs) with the full 2-D
    index list: cone 1 shearings -2..2, cone 2 shearings -2..2, low-pass
    (getShearletIdxs.m with full = 1).  Returns [Na, Nb, 11] complex."""
    if not feasible(Na, Nb):
        raise ValueError("plane too small for the 'kos'/'maxflat2' system")
    F1 = wedge_bandpass_lowpass((Na, Nb))
    F2 = wedge_bandpass_lowpass((Nb, Na))
    S = np.zeros((Na, Nb, 11), np.complex128)
    s = 0
    for cone in (1, 2):
        for sh in range(-2, 3):
            if cone == 1:
                S[:, :, s] = F1[0][:, :, -sh + 2 + 1 - 1] * np.conj(F1[1])
            else:
                S[:, :, s] = (F2[0][:, :, sh + 2 + 1 - 1] * np.conj(F2[1])).T
            s += 1
    S[:, :, 10] = F1[2]
    return S


def shearlet_idxs_3d():
    """getShearletIdxs(3, K=1, full=0): [pyramid, shearing1, shearing2] rows, LP last."""
    out = []
    for pyr in (1, 2, 3):
        for s1 in range(-2, 3):
            for s2 in range(-2, 3):
                if pyr == 1 or (pyr == 2 and abs(s2) < 2) or (abs(s1) < 2 and abs(s2) < 2):
                    out.append((pyr, s1, s2))
    out.append((0, 0, 0))
    return out


def shearlets_3d(N):
    """buildShearlet(N, 1, 0, 1, 'kos').S: [N1, N2, N3, 50] complex."""
    N1, N2, N3 = N
    P12 = shearlets_2d(N1, N2)           # plane (1, 2)
    P13 = shearlets_2d(N1, N3)           # plane (1, 3)
    P32 = shearlets_2d(N3, N2)           # plane (3, 2)
    A = P12[:, :, None, :]               # [N1, N2, 1, 11]
    B = P13[:, None, :, :]               # [N1, 1, N3, 11]
    Cc = np.transpose(P32, (1, 0, 2))[None, :, :, :]  # [1, N2, N3, 11]

    def ind(cone, sh):
        return 5 * (cone - 1) + 3 + sh - 1  # getIndFromIdxs, 0-based

    bands = []
    for pyr, s1, s2 in shearlet_idxs_3d():
        if pyr == 0:
            bands.append(A[..., 10] * Cc[..., 10])
        elif pyr == 1:
            bands.append(A[..., ind(1, s1)] * Cc[..., ind(1, s2)])
        elif pyr == 2:
            bands.append(B[..., ind(1, s1)] * Cc[..., ind(2, s2)])
        else:
            bands.append(A[..., ind(2, s1)] * B[..., ind(2, s2)])
    return np.stack([np.broadcast_to(b, (N1, N2, N3)) for b in bands], axis=-1)


# ----------------------------------------------------------------------------
# CS pieces (regularize.m:27-50, CSsolver.m:84-96, traceEstimation.m)
# ----------------------------------------------------------------------------


def fft3(x):
    return np.fft.fftn(x, axes=(0, 1, 2))


def ifft3(x):
    return np.fft.ifftn(x, axes=(0, 1, 2))


def shearlet_coefficients(x, S):
    """IFFT3(conj(S_w) FFT3(x)) for every band: [N1, N2, N3, NW]."""
    xF = fft3(x)
    return np.stack([ifft3(np.conj(S[..., w]) * xF) for w in range(S.shape[-1])], axis=-1)


def shearlet_weights(x, S, epsil=1e-2, p=1.0):
    """computeShearlet then (We + epsil)^(p - 2) (CSsolver.m:55-59, 84-96)."""
    return (np.abs(shearlet_coefficients(x, S)) + epsil) ** (p - 2)


def regularize_shearlet(x, S, We, la, p=1.0):
    """regularize.m:27-50: (la p / 2) IFFT3(sum_w S_w FFT3(We_w IFFT3(conj(S_w) FFT3(x))))."""
    xF = fft3(x)
    acc = np.zeros_like(xF)
    for w in range(S.shape[-1]):
        t = ifft3(np.conj(S[..., w]) * xF) * We[..., w]
        acc += fft3(t) * S[..., w]
    return (la * p / 2) * ifft3(acc)


def trace_samples_regularizer(probes, S, We):
    """Per probe: sum_w || sqrt(We_w) IFFT3(conj(S_w) FFT3(x)) ||^2 (traceEstimation.m:82-97)."""
    out = []
    for x in probes:
        c = shearlet_coefficients(x, S)
        out.append(float(np.sum(We * np.abs(c) ** 2)))
    return np.array(out)


def trace_samples_fidelity(probes, Smaps, Wgrid):
    """Per probe: sum_c sum_k Wgrid(k1, k2) |F2(S_c x) / sqrt(N1 N2)|^2
    (traceEstimation.m:61-79 with EH.We on the k-space grid).  A weight grid
    smaller than the maps' grid is the acquired grid of SENSE folding
    (traceEstimation.m:72-76): S_c x is folded onto it (fold.m) first."""
    from oracle.disorder_oracle import fold
    NX1, NX2 = Smaps.shape[:2]
    NY1, NY2 = Wgrid.shape
    out = []
    for x in probes:
        t = 0.0
        for c in range(Smaps.shape[3]):
            v = fold(fold(Smaps[..., c] * x, 0, NX1, NY1), 1, NX2, NY2)
            F = np.fft.fft2(v, axes=(0, 1)) / math.sqrt(NY1 * NY2)
            t += float(np.sum(Wgrid[:, :, None] * np.abs(F) ** 2))
        out.append(t)
    return np.array(out)


def normalise_probe(z):
    """x / sqrt(sum x^2 / N) (traceEstimation.m:55)."""
    return z / math.sqrt(float(np.sum(z * z)) / z.size)


def trace_estimation(sample_fn, shape, rng, tol=1e-2, BlSz=4):
    """traceEstimation.m:52-110 with real N(0,1) probes drawn from rng (one
    volume after another): blocks of BlSz probes until the standard error of
    the mean falls below tol of the mean."""
    vals = []
    while True:
        probes = [normalise_probe(rng.standard_normal(shape)) for _ in range(BlSz)]
        vals += list(sample_fn(probes))
        n = len(vals)
        m = float(np.mean(vals))
        sd = float(np.std(vals, ddof=1)) / math.sqrt(n)
        if n >= 2 and sd / (m + np.finfo(np.float64).eps) < tol:
            return m


def cs_solver(y, Smaps, samp, T, Nfull, x, We_grid, nOu=2, p=1.0, epsil=1e-2, tol=1e-2, seed=0):
    """CSsolver.m:31-100 with x given: shearlets of size(x), IRLS weights,
    lambda = (2/p) tr_fid / tr_reg, then nOu rounds of 3 CG iterations
    ('None') with A.Se = 1/(sum_c |S_c|^2 + la p/2) and the shearlet term,
    re-weighting and re-estimating tr_reg between rounds.  Probes come from
    numpy's default_rng(seed): the fidelity trace first, then the
    regulariser traces.  Returns (x, lambdas)."""
    from oracle.disorder_oracle import cg_solver
    rng = np.random.default_rng(seed)
    N = x.shape
    S = shearlets_3d(N)
    We = shearlet_weights(x, S, epsil, p)
    tr_fid = trace_estimation(lambda pr: trace_samples_fidelity(pr, Smaps, We_grid), N, rng, tol)
    tr_reg = trace_estimation(lambda pr: trace_samples_regularizer(pr, S, We), N, rng, tol)
    la = (2.0 / p) * tr_fid / (tr_reg + np.finfo(np.float64).eps)
    lambdas = [la]
    for it in range(1, nOu + 1):
        Se = 1.0 / (np.sum(np.abs(Smaps) ** 2, axis=-1) + la * p / 2)
        W = We
        x, _, _ = cg_solver(y, Smaps, samp, T, Nfull, x0=x, nIt=3, tolType="None", tol=tol, Se=Se,
                            reg=lambda v: regularize_shearlet(v, S, W, la, p))
        if it != nOu:
            We = shearlet_weights(x, S, epsil, p)
            tr_reg = trace_estimation(lambda pr: trace_samples_regularizer(pr, S, We), N, rng, tol)
            la = (2.0 / p) * tr_fid / (tr_reg + np.finfo(np.float64).eps)
            lambdas.append(la)
    return x, lambdas
