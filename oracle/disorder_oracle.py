"""FP64 CPU oracle for DISORDER's aligned-reconstruction hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or, in ``bench.py``, as the timed CPU baseline).  The product path in
``disorder_amd`` never imports it.

This is a NumPy restatement, written from the text of the reference MATLAB
sources, of the functions on the hot path.  Every function cites the reference
file:line it follows.  All arithmetic is float64 / complex128.

Array conventions (MATLAB column-major semantics, 0-based here):
  * a volume is an ndarray of shape (N1, N2, N3); dims 1-2 are phase encode,
    dim 3 is the readout (``Reconstruction/solveXT.m:51-56``);
  * a stack of per-state volumes has shape (N1, N2, N3, S): the state axis is
    the LAST axis here (MATLAB keeps it on dim 5);
  * coil maps have shape (N1, N2, N3, C);
  * compact k-space data has shape (Ns, N3, C): samples of all states are
    concatenated in state order (``Reconstruction/encode.m:146``,
    ``Methods/buildHarmonicSampling.m:36-49``);
  * motion parameters T have shape (S, 6): T[:, 0:3] translations in
    full-resolution voxels, T[:, 3:6] rotations in radians
    (``Methods/sincRigidTransform.m:33-55``).

Pinning (see tests/test_oracle_*.py): closed forms (numpy.fft against the
literal dftmtx, integer shifts against numpy.roll, pi/2 rotations against
index permutations), invariants (unitarity, inverse(forward) = identity,
adjoint dot tests, row-DFT form == FFT+mask form), central finite differences
of the forward transform for the gradient, a dense-matrix least-squares
solution for CG, and recovery of a known motion for LM.
"""
from __future__ import annotations

import itertools
import math

import numpy as np

_FFT_WORKERS = None  # None: numpy.fft; an int: scipy.fft with that many workers


def set_fft_workers(n):
    """Use scipy.fft with ``n`` threads (None restores numpy.fft).  Only the
    CPU-baseline timing uses this; results are identical to rounding."""
    global _FFT_WORKERS
    _FFT_WORKERS = n


_DENSE_DFT = False
_DFT_CACHE = {}


def set_dense_dft(flag):
    """Route every DFT along an axis through a GEMM with the single-precision
    matrix single(dftmtx(n)) (inverse conj(F)/n), as Methods/fftGPU.m:28-45
    and ifftGPU.m:24-48 apply it: the reference's own algorithm, used only to
    time it on the host.  Results equal the FFT path to single precision."""
    global _DENSE_DFT
    _DENSE_DFT = bool(flag)


def _dense(x, axis, inverse):
    n = x.shape[axis]
    F = _DFT_CACHE.get(n)
    if F is None:
        F = _DFT_CACHE[n] = dftmtx(n).astype(np.complex64)
    M = np.conj(F) if inverse else F
    y = np.tensordot(M, np.asarray(x, np.complex64), axes=(1, axis))   # cgemm on the host BLAS
    y = np.moveaxis(y, 0, axis)
    return y / np.float32(n) if inverse else y


def _fft(x, axis):
    if _DENSE_DFT:
        return _dense(x, axis, False)
    if _FFT_WORKERS:
        import scipy.fft as sf
        return sf.fft(x, axis=axis, workers=_FFT_WORKERS)
    return np.fft.fft(x, axis=axis)


def _ifft(x, axis):
    if _DENSE_DFT:
        return _dense(x, axis, True)
    if _FFT_WORKERS:
        import scipy.fft as sf
        return sf.ifft(x, axis=axis, workers=_FFT_WORKERS)
    return np.fft.ifft(x, axis=axis)


def _fft2o(x):
    if _DENSE_DFT:
        return _dense(_dense(x, 0, False), 1, False) / np.float32(math.sqrt(x.shape[0] * x.shape[1]))
    if _FFT_WORKERS:
        import scipy.fft as sf
        return sf.fft2(x, axes=(0, 1), norm="ortho", workers=_FFT_WORKERS)
    return np.fft.fft2(x, axes=(0, 1), norm="ortho")


def _ifft2o(x):
    if _DENSE_DFT:
        return _dense(_dense(x, 0, True), 1, True) * np.float32(math.sqrt(x.shape[0] * x.shape[1]))
    if _FFT_WORKERS:
        import scipy.fft as sf
        return sf.ifft2(x, axes=(0, 1), norm="ortho", workers=_FFT_WORKERS)
    return np.fft.ifft2(x, axes=(0, 1), norm="ortho")

# Rotation m (0-based) acts in the plane of axes (PER[0][m], PER[1][m]):
# per=[1 3 2;2 1 3] in Methods/generateTransformGrids.m:55-56 and
# Methods/sincRigidTransform.m:33-34.
PER = ((0, 2, 1), (1, 0, 2))
# Hessian pairs ha=[diag(1..6), nchoosek(1:6,2)] (Reconstruction/LMsolver.m:26)
HA = [(g, g) for g in range(6)] + list(itertools.combinations(range(6), 2))


# ----------------------------------------------------------------------------
# Angle helpers
# ----------------------------------------------------------------------------
def wrap_to_pi(a):
    """MATLAB Mapping Toolbox ``wrapToPi`` (call site
    Methods/precomputeFactorsSincRigidTransform.m:46): angles already in
    [-pi, pi] are unchanged; others are wrapped, positive multiples of pi map
    to +pi."""
    a = np.array(a, dtype=np.float64, copy=True)
    q = (a < -np.pi) | (a > np.pi)
    if np.any(q):
        b = a[q] + np.pi
        pos = b > 0
        b = np.mod(b, 2 * np.pi)
        b[(b == 0) & pos] = 2 * np.pi
        a[q] = b - np.pi
    return a


def wrap_to_pi_half(a):
    """Methods/wrapToPiHalf.m:9-12: mod(x,pi), then subtract pi above pi/2."""
    a = np.mod(np.asarray(a, dtype=np.float64), np.pi)
    return np.where(a > np.pi / 2, a - np.pi, a)


def restrict_transform(T):
    """Methods/restrictTransform.m:28-29 (no limits): rotations into [-pi, pi]
    by repeated +-2pi."""
    T = np.array(T, dtype=np.float64, copy=True)
    ang = T[:, 3:6]
    far = np.abs(ang) > 64 * np.pi  # bounded loop for far-out values
    ang[far] = np.fmod(ang[far], 2 * np.pi)
    while np.any(ang > np.pi):
        ang[ang > np.pi] -= 2 * np.pi
    while np.any(ang < -np.pi):
        ang[ang < -np.pi] += 2 * np.pi
    T[:, 3:6] = ang
    return T


# ----------------------------------------------------------------------------
# Grids (Methods/generateGrid.m:34, Methods/generateTransformGrids.m:48-61)
# ----------------------------------------------------------------------------
def generate_grid(N, nor, cent):
    """Methods/generateGrid.m:33-39: p_m[j] = (nor_m/N_m) * (j - cent_m),
    j = 1..N_m."""
    return [(float(nor[m]) / N[m]) * (np.arange(1, N[m] + 1, dtype=np.float64) - cent[m])
            for m in range(len(N))]


def transform_grids(Nfull, n):
    """Spatial and (ifftshifted) spectral grids of a pyramid level.

    Methods/generateTransformGrids.m:48-52 as called from
    Methods/downsampleOperators.m:62-63 (cent=ceil((n+1)/2), sh=1):
      r_d[j] = (N_d/n_d)(j - ceil((n_d+1)/2)),
      k_d    = ifftshift((2 pi/N_d)(j - ceil((n_d+1)/2))).
    """
    n = [int(v) for v in n]
    Nfull = [float(v) for v in Nfull]
    cent = [math.ceil((v + 1) / 2) for v in n]
    r = generate_grid(n, Nfull, cent)
    k = generate_grid(n, [2 * np.pi * n[d] / Nfull[d] for d in range(3)], cent)
    k = [np.fft.ifftshift(km) for km in k]
    return r, k


def _along(v, axis, ndim=4):
    shp = [1] * ndim
    shp[axis] = v.shape[0]
    return v.reshape(shp)


def _plane(coef, ra, axis_r, kb, axis_k):
    """exp-argument plane coef[s] * r(axis_r) * k(axis_k), shape (.., S)."""
    return (_along(ra, axis_r) * _along(kb, axis_k)) * coef.reshape(1, 1, 1, -1)


# ----------------------------------------------------------------------------
# Factors (Methods/precomputeFactorsSincRigidTransform.m:45-131, gr=1, sh=1)
# ----------------------------------------------------------------------------
class Factors:
    """Per-state transform factors ``et``/``etg`` for the grouped (gr=1),
    pre-shifted (sh=1) case used by the reconstruction path.

    e1[d]  : (.., S) arrays along axis d        exp(t_d k_d)          (:115)
    e2[m]  : plane over (a=PER0[m], b=PER1[m])  exp(tan2j_m rk1_m)    (:79)
    e3[m]  : same plane                          exp(sinj_m  rk2_m)    (:80)
    flip   : (S,3) bool, |wrapToPi(theta)| > pi/2                      (:48-50)
    cgrid  : flip centres n/2+1                                        (:31,52)
    etg1[d]: -i k_d                                                    (:123)
    etg2[m]: i (1+tan^2(th/2))/2 rk1_m e2_m                            (:89,99)
    etg3[m]: -i cos(th) rk2_m e3_m                                     (:90,99)
    """

    def __init__(self, T, n, Nfull, di, cg=False):
        T = np.atleast_2d(np.asarray(T, dtype=np.float64))
        self.S = T.shape[0]
        self.n = [int(v) for v in n]
        self.di = di
        r, k = transform_grids(Nfull, n)
        self.r, self.k = r, k
        sgn = (-1.0) ** di
        t = sgn * 1j * T[:, 0:3]
        th = wrap_to_pi(T[:, 3:6])
        self.flip = np.abs(th) > np.pi / 2
        self.cgrid = [nv / 2 + 1 for nv in self.n]
        th = wrap_to_pi_half(th)
        tan2 = np.tan(th / 2)
        tan2j = ((-1.0) ** (di - 1)) * 1j * tan2
        sinj = sgn * 1j * np.sin(th)
        self.e2, self.e3 = [], []
        for m in range(3):
            a, b = PER[0][m], PER[1][m]
            # rk{1}{m} = r_{per(2,m)} .* k_{per(1,m)};  rk{2}{m} = r_{per(1,m)} .* k_{per(2,m)}
            self.e2.append(np.exp(_plane(tan2j[:, m], r[b], b, k[a], a)))
            self.e3.append(np.exp(_plane(sinj[:, m], r[a], a, k[b], b)))
        self.e1 = [np.exp(_along(k[d], d) * t[:, d].reshape(1, 1, 1, -1)) for d in range(3)]
        if cg:
            tc = (1 + tan2 * tan2) / 2
            co = np.cos(th)
            self.etg1 = [-1j * _along(k[d], d) for d in range(3)]
            self.etg2, self.etg3 = [], []
            for m in range(3):
                a, b = PER[0][m], PER[1][m]
                self.etg2.append(1j * _plane(tc[:, m], r[b], b, k[a], a) * self.e2[m])
                self.etg3.append(-1j * _plane(co[:, m], r[a], a, k[b], b) * self.e3[m])


# ----------------------------------------------------------------------------
# DFT helpers (Methods/fftGPU.m:37-42, ifftGPU.m:31-39, build1DFTM.m:21-52)
# ----------------------------------------------------------------------------
def dftmtx(n):
    """Methods/build1DFTM.m:21 -- dftmtx(n)(j,k) = exp(-2 pi i (j-1)(k-1)/n)."""
    j = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(j, j) / n)


def fft_dim(x, m):
    """fftGPU(x,m): non-unitary forward DFT along axis m."""
    return _fft(x, m)


def ifft_dim(x, m):
    """ifftGPU(x,m): conj(F)/n along axis m."""
    return _ifft(x, m)


# ----------------------------------------------------------------------------
# Shifting / flipping (Methods/shifting.m:25-48, Methods/flipping.m:11-22)
# ----------------------------------------------------------------------------
def shifting(x, H):
    """Methods/shifting.m:25-48 with f=0, zp=0.  H maps axis -> scalar shift.
    Integer shifts are circular index shifts (:29-33); others are Fourier
    phase ramps exp(-i h k) with k = ifftshift((2pi/n)(j-ceil((n+1)/2)))
    (:39-45)."""
    for m, h in H.items():
        n = x.shape[m]
        if float(h) == math.floor(h):
            if n != 1:
                x = np.roll(x, int(h), axis=m)
        else:
            kg = generate_grid([n], [2 * np.pi], [math.ceil((n + 1) / 2)])[0]
            kg = -1j * np.fft.ifftshift(kg)
            ph = np.exp(h * kg)
            x = ifft_dim(fft_dim(x, m) * _along(ph, m, x.ndim), m)
    return x


def flipping(x, centers):
    """Methods/flipping.m:11-22: shift by -c+1/2, flip, shift by c-1/2 on the
    axes listed in ``centers`` (axis -> centre)."""
    H = {a: -c + 0.5 for a, c in centers.items()}
    x = shifting(x, H)
    for a in centers:
        x = np.flip(x, axis=a)
    H = {a: -h for a, h in H.items()}
    return shifting(x, H)


def _flip_states(x, f, m):
    """x.*(1-et5)+flipping(x,et4).*et5 (Methods/sincRigidTransform.m:43,109)."""
    fl = f.flip[:, m]
    if not np.any(fl):
        return x
    a, b = PER[0][m], PER[1][m]
    if x.shape[3] != f.S:
        x = np.broadcast_to(x, x.shape[:3] + (f.S,)).copy()
    xf = flipping(x, {a: f.cgrid[a], b: f.cgrid[b]})
    w = fl.astype(np.float64).reshape(1, 1, 1, -1)
    return x * (1 - w) + xf * w


# ----------------------------------------------------------------------------
# Sinc rigid transform (Methods/sincRigidTransform.m:38-114)
# ----------------------------------------------------------------------------
def sinc_rigid_transform(x, f: Factors, di, sumR=True, keep=False):
    """Methods/sincRigidTransform.m:38-114 (useFull=1, gr=1 cell factors).

    di=1: x is (N1,N2,N3) or (N1,N2,N3,1|S); returns (N1,N2,N3,S) and, with
    keep, the intermediates xB[0..4] (``xB{1..5}``).
    di=0: x is (N1,N2,N3,S); returns the state sum (N1,N2,N3) when sumR.
    """
    x = np.asarray(x, dtype=np.complex128)
    if x.ndim == 3:
        x = x[..., None]
    xB = [None] * 5
    if di == 1:
        xB[4] = x
        for m in range(3):
            a, b = PER[0][m], PER[1][m]
            x = _flip_states(x, f, m)
            x = fft_dim(x, a)
            xB[m] = x
            x = ifft_dim(x * f.e2[m], a)
            x = ifft_dim(fft_dim(x, b) * f.e3[m], b)
            x = fft_dim(x, a) * f.e2[m]
            if m != 2:
                x = ifft_dim(x, a)
        for d in (0, 2):  # dim 2 (axis 1) is already in k-space (:66-68)
            x = fft_dim(x, d)
        xB[3] = x
        for d in range(3):
            x = x * f.e1[d]
        for d in (2, 1, 0):
            x = ifft_dim(x, d)
        if x.shape[3] != f.S:
            x = np.broadcast_to(x, x.shape[:3] + (f.S,)).copy()
    else:
        for d in range(3):  # (:86-90)
            x = fft_dim(x, d) * f.e1[d]
            if d != 1:
                x = ifft_dim(x, d)
        for m in (2, 1, 0):  # (:99-110)
            a, b = PER[0][m], PER[1][m]
            if m != 2:
                x = fft_dim(x, a)
            x = ifft_dim(x * f.e2[m], a)
            x = ifft_dim(fft_dim(x, b) * f.e3[m], b)
            x = ifft_dim(fft_dim(x, a) * f.e2[m], a)
            x = _flip_states(x, f, m)
        if sumR:
            x = x.sum(axis=3)
    return (x, xB) if keep else x


def rigid_forward(x, T, n, Nfull):
    return sinc_rigid_transform(x, Factors(T, n, Nfull, 1), 1)


def rigid_inverse(x, T, n, Nfull, sumR=True):
    return sinc_rigid_transform(x, Factors(T, n, Nfull, 0), 0, sumR=sumR)


# ----------------------------------------------------------------------------
# Gradient (Methods/sincRigidTransformGradient.m:28-147)
# ----------------------------------------------------------------------------
def sinc_rigid_transform_gradient(xB, f: Factors):
    """d(T x)/dT_g, g = 0..5 (Methods/sincRigidTransformGradient.m:28-147).
    ``f`` must be built with cg=True, di=1; ``xB`` from sinc_rigid_transform
    (keep=True).  Returns a list of 6 arrays (N1,N2,N3,S).  The resampling
    quirk at :149 is a no-op because Npad == N here."""
    G = [None] * 6
    S = f.S

    def bc(v):
        return v if v.shape[3] == S else np.broadcast_to(v, v.shape[:3] + (S,)).copy()

    for m in range(3):  # translations (:29-36)
        y = xB[3] * f.etg1[m]
        for d in range(3):
            y = y * f.e1[d]
        for d in range(3):
            y = ifft_dim(y, d)
        G[m] = bc(y)

    def rot_chain(m):
        """Product rule through the three shears of rotation m (:39-52)."""
        a, b = PER[0][m], PER[1][m]
        x1 = xB[m] * f.e2[m]
        x2 = xB[m] * f.etg2[m]
        x1 = fft_dim(ifft_dim(x1, a), b)
        x2 = fft_dim(ifft_dim(x2, a), b)
        x2 = x1 * f.etg3[m] + x2 * f.e3[m]
        x1 = x1 * f.e3[m]
        x1 = fft_dim(ifft_dim(x1, b), a)
        x2 = fft_dim(ifft_dim(x2, b), a)
        return x1 * f.etg2[m] + x2 * f.e2[m]

    def later_rotation(x, m):
        """Rotation m applied to an image-space volume (:54-64, :66-75)."""
        a, b = PER[0][m], PER[1][m]
        x = _flip_states(x, f, m)
        x = fft_dim(x, a) * f.e2[m]
        x = ifft_dim(x, a)
        x = ifft_dim(fft_dim(x, b) * f.e3[m], b)
        x = fft_dim(x, a) * f.e2[m]
        if m != 2:
            x = ifft_dim(x, a)
        return x

    def finish(x):
        """Translation stage (:77-85)."""
        for d in (0, 2):
            x = fft_dim(x, d)
        for d in range(3):
            x = x * f.e1[d]
        for d in (2, 1, 0):
            x = ifft_dim(x, d)
        return x

    for m in range(3):
        x = rot_chain(m)
        a = PER[0][m]
        if m != 2:
            x = ifft_dim(x, a)
            for mm in range(m + 1, 3):
                x = later_rotation(x, mm)
        G[3 + m] = bc(finish(x))
    return G


# ----------------------------------------------------------------------------
# Sampling (Methods/buildHarmonicSampling.m:36-49)
# ----------------------------------------------------------------------------
class Sampling:
    """Per-state sample lists.  k1/k2 are 0-based FFT-order indices into the
    (n1, n2) phase-encode grid; samples are concatenated in state order;
    ``off`` holds the state offsets (``E.nSt``)."""

    def __init__(self, k1, k2, off, n1, n2, We=None):
        self.k1 = np.asarray(k1, dtype=np.int64)
        self.k2 = np.asarray(k2, dtype=np.int64)
        self.off = np.asarray(off, dtype=np.int64)
        self.n1, self.n2 = int(n1), int(n2)
        self.We = None if We is None else np.asarray(We, dtype=np.float64)

    @property
    def S(self):
        return len(self.off) - 1

    @property
    def Ns(self):
        return int(self.off[-1])

    def state(self, s):
        sl = slice(int(self.off[s]), int(self.off[s + 1]))
        return sl, self.k1[sl], self.k2[sl]

    def state_of_sample(self):
        """``E.mSt`` (0-based)."""
        return np.repeat(np.arange(self.S), np.diff(self.off))


def build_harmonic_sampling(H, nstates):
    """Methods/buildHarmonicSampling.m:36-49 (integer part).

    H: (n1, n2) or (n1, n2, repeats) int array; H[k1,k2(,r)] = state
    (1..nstates) that sampled that phase-encode position (in repeat r), 0 if
    unsampled.  For s = 1..nstates+1 the samples of state mod(s, nstates+1)
    are taken in MATLAB linear (column-major) order over all of H's
    dimensions (:27-28, :36-38), so the last group is the unsampled
    positions (:37).  Returns (k1, k2, off, idx) with off of length
    nstates+2 (``E.nSt`` with the outer group appended, :49-50) and idx the
    column-major linear indices; the repeat of a sample is idx // (n1 n2),
    which is also its row in the data permuted to [n1, n2, repeats, ...]
    (:29-30).
    """
    n1, n2 = H.shape[:2]
    Hl = H.reshape(-1, order="F")
    k1s, k2s, idx, off = [], [], [], [0]
    for s in range(1, nstates + 2):
        ind = np.nonzero(Hl == (s % (nstates + 1)))[0]
        idx.append(ind)
        k1s.append(ind % n1)
        k2s.append((ind // n1) % n2)
        off.append(off[-1] + len(ind))
    return (np.concatenate(k1s), np.concatenate(k2s), np.asarray(off, np.int64),
            np.concatenate(idx))


# ----------------------------------------------------------------------------
# SENSE folding (Methods/fold.m, ifold.m, buildFoldM.m)
# ----------------------------------------------------------------------------
def _fold_consts(NS, NY):
    odd = 2 * math.ceil((math.ceil(NS / NY) - 1) / 2) + 1      # oddFactSENSE
    over = (NS - NY) / 2
    return odd, (math.floor(over), math.ceil(over))


def fold(x, m, NS, NY):
    """Methods/fold.m:30-110 with A empty and fl = 0: the NS samples of axis
    m aliased onto NY.  Non-multiples: the central window
    x(FOV(2)+1:NS-FOV(1)) plus the windows shifted by +-s NY (:52-68) and the
    two partial wraps at the ends (:85-100); multiples: reshape to
    [NY, NS/NY], sum, circshift by -FOV(2) (:104-109)."""
    if NS == NY:
        return x
    x = np.moveaxis(np.asarray(x), m, 0)
    odd, FOV = _fold_consts(NS, NY)
    oFRed = (odd - 3) // 2
    oFRedT = oFRed * NY
    if NS % NY != 0:
        xb = x[FOV[1]:NS - FOV[0]].copy()
        for s_ in range(1, oFRed + 1):
            xb += x[FOV[1] - s_ * NY:NS - FOV[0] - s_ * NY] + x[FOV[1] + s_ * NY:NS - FOV[0] + s_ * NY]
        xb[0:FOV[0] - oFRedT] += x[NS - FOV[0] + oFRedT:NS]
        xb[NY - FOV[1] + oFRedT:NY] += x[0:FOV[1] - oFRedT]
    else:
        q = NS // NY
        xb = x.reshape((q, NY) + x.shape[1:]).sum(axis=0)     # element i + NY k -> row i
        xb = np.roll(xb, -FOV[1], axis=0)
    return np.moveaxis(xb, 0, m)


def ifold(x, m, NS, NY):
    """Methods/ifold.m:28-72 with A empty and fl = 0: unfolding NY -> NS,
    indUnf = [1+iFOV(1):NY, repmat(1:NY, oFRed), 1:NY-iFOV(2)] (:46) or, for
    multiples, circshift by -iFOV(1) and repmat (:66-70)."""
    if NS == NY:
        return x
    x = np.moveaxis(np.asarray(x), m, 0)
    odd, _ = _fold_consts(NS, NY)
    oFRed = odd - 2
    disc = (odd * NY - NS) / 2
    iFOV = (math.floor(disc), math.ceil(disc))
    if NS % NY != 0:
        ind = np.r_[np.arange(iFOV[0], NY), np.tile(np.arange(NY), oFRed), np.arange(0, NY - iFOV[1])]
        out = x[ind]
    else:
        out = np.tile(np.roll(x, -iFOV[0], axis=0), (NS // NY,) + (1,) * (x.ndim - 1))
    return np.moveaxis(out, 0, m)


def build_fold_matrix(NS, NY):
    """Methods/buildFoldM.m:37-48 (emre = 0): the [NY x NS] 0/1 matrix A
    with fold(x) = A x; the central run of 2 oFRed + 1 windows (:44) and the
    two end wraps (:45), set through MATLAB column-major linear indices."""
    A = np.zeros(NY * NS)
    if NS == NY:
        return np.eye(NY)
    odd, FOV = _fold_consts(NS, NY)
    oFRed = (odd - 3) // 2
    oFRedT = oFRed * NY
    rows = np.tile(np.arange(1, NY + 1), 2 * oFRed + 1)
    cols = np.arange(FOV[1] - oFRed * NY, NS - FOV[0] + oFRed * NY)        # 0-based column (MATLAB col - 1)
    A[rows + cols * NY - 1] = 1
    r2 = np.r_[np.arange(1, FOV[0] - oFRedT + 1), np.arange(NY - FOV[1] + 1 + oFRedT, NY + 1)]
    c2 = np.r_[np.arange(NS - FOV[0] + 1 + oFRedT, NS + 1), np.arange(1, FOV[1] - oFRedT + 1)]
    A[r2 + (c2 - 1) * NY - 1] = 1
    return A.reshape((NY, NS), order="F")


def _fold12(v, n, ny):
    for m in range(2):
        v = fold(v, m, n[m], ny[m])
    return v


def _ifold12(v, n, ny):
    for m in range(2):
        v = ifold(v, m, n[m], ny[m])
    return v


# ----------------------------------------------------------------------------
# Operators (Reconstruction/encode.m, decode.m, encodedecode.m)
# ----------------------------------------------------------------------------
def _transform_states(x, T, n, Nfull, nmotion, S):
    if np.any(T[:nmotion] != 0):  # encode.m:61
        xt = rigid_forward(x, T[:nmotion], n, Nfull)
    else:
        xt = np.broadcast_to(x[..., None], x.shape + (nmotion,))
    if S > nmotion:  # outer segments use the untransformed x (encode.m:76-79)
        xt = np.concatenate([xt, np.broadcast_to(x[..., None], x.shape + (S - nmotion,))], axis=3)
    return xt


def _inverse_states(v, T, n, Nfull, nmotion):
    if np.any(T[:nmotion] != 0):  # decode.m:97-111
        out = rigid_inverse(v[..., :nmotion], T[:nmotion], n, Nfull, sumR=True)
    else:
        out = v[..., :nmotion].sum(axis=3)
    return out + v[..., nmotion:].sum(axis=3)


def encode(x, Smaps, samp: Sampling, T, Nfull, nmotion=None):
    """Reconstruction/encode.m:55-146 (Tr, Sf, Fs fields): y = E x, shape
    (Ns, N3, C).  Per state: T_s x (:59-80), coil multiply (:93-96), then the
    unitary 2D DFT over dims 1-2 evaluated at the state's samples
    (:120-136, Fs rows from buildStandardDFTM(NY,1))."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    n = x.shape
    nmotion = T.shape[0] if nmotion is None else nmotion
    xt = _transform_states(x, T, n, Nfull, nmotion, samp.S)
    C = Smaps.shape[3]
    y = np.zeros((samp.Ns, n[2], C), np.complex128)
    ny = (samp.n1, samp.n2)
    for s in range(samp.S):
        sl, k1, k2 = samp.state(s)
        if len(k1) == 0:
            continue
        for c in range(C):
            X = _fft2o(_fold12(Smaps[..., c] * xt[..., s], n, ny))   # SENSE folding, encode.m:100-104
            y[sl, :, c] = X[k1, k2, :]
    return y


def encode_rowdft(x, Smaps, samp: Sampling, T, Nfull, nmotion=None):
    """Literal row-DFT form of Reconstruction/encode.m:120-136: aplGPU with
    the selected rows of F/sqrt(n1), then sum(xQ .* F2rows, 2)."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    n = x.shape
    nmotion = T.shape[0] if nmotion is None else nmotion
    xt = _transform_states(x, T, n, Nfull, nmotion, samp.S)
    C = Smaps.shape[3]
    F1 = dftmtx(n[0]) / np.sqrt(n[0])
    F2 = dftmtx(n[1]) / np.sqrt(n[1])
    y = np.zeros((samp.Ns, n[2], C), np.complex128)
    for s in range(samp.S):
        sl, k1, k2 = samp.state(s)
        for c in range(C):
            xs = Smaps[..., c] * xt[..., s]
            xq = np.tensordot(F1[k1, :], xs, axes=(1, 0))          # (Nec, N2, N3)
            y[sl, :, c] = np.sum(xq * F2[k2, :][:, :, None], axis=1)
    return y


def decode(y, Smaps, samp: Sampling, T, Nfull, nmotion=None):
    """Reconstruction/decode.m:25-120: x = E^H (We .* y).  Per state: weights
    (:46), conj(F2 rows) and F1 rows' adjoint (:51-52), conj(S) and coil sum
    (:82-84), inverse transform (di=0) and state sum (:96-112)."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    ny1, ny2 = samp.n1, samp.n2
    n1, n2 = Smaps.shape[:2]
    n3 = y.shape[1]
    C = Smaps.shape[3]
    S = samp.S
    nmotion = S if nmotion is None else nmotion
    v = np.zeros((n1, n2, n3, S), np.complex128)
    for s in range(S):
        sl, k1, k2 = samp.state(s)
        ys = y[sl]
        if samp.We is not None:
            ys = ys * samp.We[sl][:, None, None]
        for c in range(C):
            K = np.zeros((ny1, ny2, n3), np.complex128)
            np.add.at(K, (k1, k2), ys[:, :, c])
            # SENSE unfolding (decode.m:73-76)
            v[..., s] += np.conj(Smaps[..., c]) * _ifold12(_ifft2o(K), (n1, n2), (ny1, ny2))
    return _inverse_states(v, T, (n1, n2, n3), Nfull, nmotion)


def encodedecode(x, Smaps, samp: Sampling, T, Nfull, nmotion=None):
    """Reconstruction/encodedecode.m:56-211: E^H W E x (weights applied once,
    :133)."""
    return decode(encode(x, Smaps, samp, T, Nfull, nmotion), Smaps, samp, T, Nfull, nmotion)


# ----------------------------------------------------------------------------
# Energy, CG (Reconstruction/computeEnergy.m, CGsolver.m)
# ----------------------------------------------------------------------------
def compute_energy(y, x, Smaps, samp, T, Nfull, nmotion=None, per_sample=False, Mbe=None):
    """Reconstruction/computeEnergy.m:34-57 (no=2): residual E x - y, optional
    readout weights sqrt(Mbe) (:39); per-sample sums over readout and coils
    with avdi (:42), else the total energy."""
    res = encode(x, Smaps, samp, T, Nfull, nmotion) - y
    a2 = np.abs(res) ** 2
    if Mbe is not None:
        a2 = a2 * np.asarray(Mbe, np.float64)[None, :, None]
    if per_sample:
        return a2.sum(axis=(1, 2))
    return float(a2.sum())


def cg_solver(y, Smaps, samp, T, Nfull, x0=None, nIt=300, tolType="Energy", tol=1e-5,
              Se=None, Ma=None, lam=None, nmotion=None, reg=None):
    """Reconstruction/CGsolver.m:45-109 with A.Se (precondition.m:19-23),
    C.Ma (constrain.m:12), an optional Tikhonov term R.Ti.la (regularize.m:24)
    and an optional further regulariser reg(v) added to R x (the shearlet term
    of regularize.m:27-50 for CSsolver).

    Returns (x, En, nn) where En holds [fidelity; regulariser] energies of the
    odd iterations for tolType 'Energy' (computeEnergy.m:72-80)."""
    def C_(v):
        return v if Ma is None else v * Ma

    def P_(v):
        return v if Se is None else v * Se

    def R_(v):
        r = 0 if lam is None else lam * v
        return r if reg is None else r + reg(v)

    def EHE(v):
        return C_(encodedecode(v, Smaps, samp, T, Nfull, nmotion) + R_(v))

    def energy(v):
        fid = compute_energy(y, v, Smaps, samp, T, Nfull, nmotion)
        reg = 0.0 if lam is None else float(abs(np.sum(lam * v * np.conj(v))))
        return np.array([fid, reg])

    tolTyp = {"RelativeResidual2Error": 1, "Energy": 6, "None": 7}[tolType]
    NitTest = 2
    b = C_(decode(y, Smaps, samp, T, Nfull, nmotion))
    if x0 is None:
        x = np.zeros_like(b)
        ehe = np.zeros_like(b)
    else:
        x = np.array(x0, np.complex128, copy=True)
        ehe = EHE(x)
    bnorm2 = None
    if tolTyp not in (6, 7):
        zb = np.conj(P_(b)) * b
        bnorm2 = float(np.sum(np.abs(zb)))
    r = b - ehe
    z = P_(r)
    p = z.copy()
    rsold = np.sum(np.conj(z) * r)
    if abs(rsold) < 1e-6:  # :70-75
        return x, np.zeros((2, 0)), 0
    err = np.inf
    En = {}
    n = 0
    for n in range(1, nIt + 1):
        ehe = EHE(p)
        g = rsold / np.sum(np.conj(p) * ehe)
        x = x + g * p
        if n % NitTest == 1 and tolTyp == 6:
            En[n] = energy(x)
        r = r - g * ehe
        z = P_(r)
        zr = np.conj(z) * r
        rs = np.sum(zr)
        d = rs / (rsold + np.finfo(np.float64).eps)
        p = z + d * p
        rsold = rs
        if tolTyp == 1:
            err = math.sqrt(float(np.sum(np.abs(zr))) / bnorm2)
        elif tolTyp == 6 and n % NitTest == 1 and n > 1:
            e0, e1 = En[n - NitTest].sum(), En[n].sum()
            err = (e0 - e1) / e0
        if err < tol or abs(rs) < 1e-6:
            break
    Ek = sorted(En)
    return x, (np.stack([En[k] for k in Ek], axis=1) if Ek else np.zeros((2, 0))), n


# ----------------------------------------------------------------------------
# LM (Reconstruction/LMsolver.m:43-225)
# ----------------------------------------------------------------------------
def readout_roi(n3, fat_shift_dir="H", redFOV=1.0 / 3.0):
    """E.nF of Reconstruction/solveXT.m:345 as a 0-based half-open range
    [lo, hi): 'F' keeps 1:floor((1-redFOV) n3), 'H' keeps
    floor(1+redFOV n3):n3 (1-based, inclusive), anything else every plane."""
    if fat_shift_dir == "F":
        return 0, int(math.floor((1.0 - redFOV) * n3))
    if fat_shift_dir == "H":
        return int(math.floor(1.0 + redFOV * n3)) - 1, int(n3)
    return 0, int(n3)


def _crop_readout(a, nF, axis):
    """dynInd(a, E.nF, axis) for a range nF = (lo, hi) (LMsolver.m:33,68,98)."""
    if nF is None:
        return a
    idx = [slice(None)] * a.ndim
    idx[axis] = slice(nF[0], nF[1])
    return a[tuple(idx)]


def fractional_filter(N, sp, order):
    """Methods/buildFilter.m 'FractionalFiniteDiscreteIso' on the grid N
    (restated literally over the whole grid): kGrid = generateGrid(N, gpu,
    2 pi i ./ sp, ceil((N+1)/2)) (:30), x = sum over dimensions with N > 1
    of |kGrid|^2 (:59-61), sqrt (:62), clipped at pi (:64), 2 |sin(x/2)|
    (:65-66), x.^order (:68), ifftshift (:120), then x(1) = max(x(2:end))
    (:134: the zero frequency takes the largest value of the rest)."""
    N = [int(v) for v in N]
    x = np.zeros(tuple(N))
    for m, n in enumerate(N):
        if n > 1:
            g = generate_grid([n], [2 * np.pi / sp[m]], [math.ceil((n + 1) / 2)])[0]  # |2 pi i / sp| scale
            x = x + np.abs(g.reshape([-1 if d == m else 1 for d in range(len(N))])) ** 2
    x = np.sqrt(x)
    x[x > np.pi] = np.pi
    x = 2 * np.abs(np.sin(x / 2))
    with np.errstate(divide="ignore"):
        x = x ** order
    x = np.fft.ifftshift(x)
    flat = x.reshape(-1)
    flat[0] = np.max(flat[1:])
    return flat.reshape(x.shape)


def fractional_samples(F3, samp):
    """GSt of Methods/buildHarmonicSampling.m:35,42: the filter rows at the
    samples' phase-encode positions, [Ns, nF]."""
    return F3[np.asarray(samp.k1), np.asarray(samp.k2), :]


def _weighted(a, Fd_rows):
    """fftGPU(a, 3) .* E.Fd (Reconstruction/LMsolver.m:89-90, 102-103):
    unnormalised DFT along the readout of [samples, nF, C] rows, times the
    per-sample weights [samples, nF]."""
    if Fd_rows is None:
        return a
    return fft_dim(a, 1) * Fd_rows[:, :, None]


def lm_partials(y, x, Smaps, samp, T, Nfull, states=None, nF=None, Fd=None):
    """Per-state residual energy, gradient dG and Gauss-Newton Hessian dH
    (Reconstruction/LMsolver.m:52-122), all states transformed (NMs = S).
    nF: readout ROI (lo, hi) or None: the maps and data are cropped to it
    (:33), and so are the transformed image (:68) and its derivatives (:98)
    before encoding; energies are still divided by numel(y) of the uncropped
    data (:32).  Fd: E.Fd [Ns, nF] or None: residual and Jacobians pass
    through the readout DFT and the weights first (:88-91, :101-104).
    Returns EnPrev (S,), dG (S,6), dH (S,21)."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    S = T.shape[0]
    states = range(S) if states is None else states
    NElY = y.size
    Smaps = _crop_readout(Smaps, nF, 2)
    y = _crop_readout(y, nF, 1)
    n = x.shape
    En = np.zeros(S)
    dG = np.zeros((S, 6))
    dH = np.zeros((S, 21))
    for s in states:
        f = Factors(T[s:s + 1], n, Nfull, 1, cg=True)
        xt, xB = sinc_rigid_transform(x, f, 1, keep=True)
        G = sinc_rigid_transform_gradient(xB, f)
        sub = _one_state(samp, s)
        Tz = np.zeros((1, 6))
        rows = samp.state(s)[0]
        Fr = None if Fd is None else Fd[rows]
        r = _weighted(encode(_crop_readout(xt[..., 0], nF, 2), Smaps, sub, Tz, Nfull) - y[rows], Fr)
        En[s] = float(np.sum(np.abs(r) ** 2)) / NElY
        J = [_weighted(encode(_crop_readout(G[g][..., 0], nF, 2), Smaps, sub, Tz, Nfull), Fr) for g in range(6)]
        for g in range(6):
            dG[s, g] = float(np.sum(np.real(J[g] * np.conj(r))))
        for h, (i, j) in enumerate(HA):
            dH[s, h] = float(np.sum(np.real(J[i] * np.conj(J[j]))))
    return En, dG, dH


def _one_state(samp, s):
    sl, k1, k2 = samp.state(s)
    return Sampling(k1, k2, [0, len(k1)], samp.n1, samp.n2,
                    None if samp.We is None else samp.We[sl])


def lm_state_energy(y, x, Smaps, samp, T, Nfull, states, nF=None, Fd=None):
    """Trial energies of LMsolver.m:151-192: transform (:166), readout crop
    (:167), encode minus the cropped data (:171), with E.Fd the readout DFT
    and weights (:186-189), over numel(y) (:32)."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    NElY = y.size
    Sc, yc = _crop_readout(Smaps, nF, 2), _crop_readout(y, nF, 1)
    out = {}
    for s in states:
        sub = _one_state(samp, s)
        xt = rigid_forward(x, T[s:s + 1], x.shape, Nfull)[..., 0] if np.any(T[s] != 0) else x
        rows = samp.state(s)[0]
        r = encode(_crop_readout(xt, nF, 2), Sc, sub, np.zeros((1, 6)), Nfull) - yc[rows]
        r = _weighted(r, None if Fd is None else Fd[rows])
        out[s] = float(np.sum(np.abs(r) ** 2)) / NElY
    return out


def lm_step(dH, dG, w, winit):
    """Damped 6x6 FP64 solve (Reconstruction/LMsolver.m:132-141):
    MH_ii=(1+w)H_ii, MH_ij=H_ij, step = -(winit/w) MH\\g."""
    MH = np.zeros((6, 6))
    for h, (i, j) in enumerate(HA):
        if i == j:
            MH[i, i] = (1 + w) * dH[h]
        else:
            MH[i, j] = MH[j, i] = dH[h]
    try:
        step = -(winit / w) * np.linalg.solve(MH, dG)
    except np.linalg.LinAlgError:
        step = np.full(6, np.nan)
    # a singular system (Inf/NaN from mldivide in the reference) takes no step
    return step if np.all(np.isfinite(step)) else np.zeros(6)


def lm_solver(y, x, Smaps, samp, T, Nfull, w, cT, winit, tL, log=None, nF=None, Fd=None):
    """Reconstruction/LMsolver.m:43-225 for one outer call, with the readout
    ROI E.nF = nF (lo, hi) or None and the fractional finite-difference
    weights E.Fd = Fd [Ns, nF] or None.  Returns the new (T, w, cT, Enmin)."""
    T = np.array(np.atleast_2d(T), np.float64, copy=True)
    w = np.array(w, np.float64, copy=True)
    cT = np.array(cT, bool, copy=True)
    S = T.shape[0]
    est = [s for s in range(S) if not cT[s]]
    EnPrev, dG, dH = lm_partials(y, x, Smaps, samp, T, Nfull, est, nF=nF, Fd=Fd)
    En = EnPrev.copy()
    dGEff = np.zeros((S, 6))
    flagw = np.zeros(S, bool)
    while True:
        for s in range(S):
            if not cT[s] and w[s] < 1e10:
                dGEff[s] = lm_step(dH[s], dG[s], w[s], winit)
        Tup = restrict_transform(T + dGEff)
        cfull = cT | flagw
        todo = [s for s in range(S) if not cfull[s]]
        for s, e in lm_state_energy(y, x, Smaps, samp, Tup, Nfull, todo, nF=nF, Fd=Fd).items():
            En[s] = e
        big = w > 1e10
        En[big] = EnPrev[big]
        flagw[En <= EnPrev] = True
        if log is not None:
            log.append((EnPrev.sum(), En.sum()))
        if not np.all(flagw):
            rej = (En > EnPrev) & ~cT
            w[rej] *= 2.0
        else:
            w[~cT] /= 1.2
            w[w < 1e-8] *= 1.2
            T = Tup
            tra = np.max(np.abs(dGEff[:, 0:3]), axis=1)
            rot = np.max(np.abs(dGEff[:, 3:6] * 180 / np.pi), axis=1)
            cT[(tra < tL[0]) & (rot < tL[1])] = True
            break
    return T, w, cT, np.minimum(EnPrev, En)


# ----------------------------------------------------------------------------
# Coil compression (Methods/compressCoils.m:75-118)
# ----------------------------------------------------------------------------
def compress_coils(Smaps, perc, y=None, ncomp=None):
    """Methods/compressCoils.m:75-118: Shat = S/(sum|S|^2+1e-9); P = sum_vox
    Shat S^H (C x C); U from svd(P); keep Cc columns by energy fraction perc
    (or ncomp); S <- U(:,1:Cc)^H S, y <- U(:,1:Cc)^H y along coils."""
    C = Smaps.shape[-1]
    Sv = Smaps.reshape(-1, C)
    nrm = np.sum(np.abs(Sv) ** 2, axis=1, keepdims=True) + 1e-9
    Sh = Sv / nrm
    P = Sh.T @ np.conj(Sv)          # P(i,j) = sum_v Shat_i conj(S_j)
    U, D, _ = np.linalg.svd(P)
    if ncomp is None:
        cs = np.cumsum(D) / np.sum(D)
        ncomp = int(np.argmax(cs >= perc) + 1)
    A = np.conj(U[:, :ncomp]).T     # U(:,1:Cc)'
    Sc = (Sv @ A.T).reshape(Smaps.shape[:-1] + (ncomp,))
    yc = None if y is None else (y.reshape(-1, C) @ A.T).reshape(y.shape[:-1] + (ncomp,))
    return Sc, yc, P, A


# ----------------------------------------------------------------------------
# Pyramid helpers (Methods/resampling.m:52-89, Methods/downsampleOperators.m:34)
# ----------------------------------------------------------------------------
def resample_fourier(x, nres):
    """Methods/resampling.m:45-80 (fo=0, mirror=0): per axis, fft/N, fftshift,
    crop/zero-pad centred at ceil((Nmax+1)/2), ifftshift, ifft*Nres."""
    x = np.asarray(x, np.complex128)
    for m, nr in enumerate(nres):
        no = x.shape[m]
        if no == nr:
            continue
        nmin, nmax = min(no, nr), max(no, nr)
        zero = math.ceil((nmax + 1) / 2)
        orig = zero - math.ceil((nmin - 1) / 2)
        fina = zero + math.floor((nmin - 1) / 2)
        X = np.fft.fftshift(np.fft.fft(x, axis=m) / no, axes=m)
        shp = list(x.shape)
        shp[m] = nr
        if no < nr:
            Y = np.zeros(shp, np.complex128)
            idx = [slice(None)] * x.ndim
            idx[m] = slice(orig - 1, fina)
            Y[tuple(idx)] = X
        else:
            idx = [slice(None)] * x.ndim
            idx[m] = slice(orig - 1, fina)
            Y = X[tuple(idx)]
        x = np.fft.ifft(np.fft.ifftshift(Y, axes=m), axis=m) * nr
    return x


def resample_kspace(A, nres):
    """Methods/resampling.m with fo=1 (k-space input, DC first): per axis
    fftshift, crop or zero-pad about ceil((Nmax+1)/2), ifftshift; no FFT."""
    A = np.asarray(A)
    for m, nr in enumerate(nres):
        no = A.shape[m]
        if no == nr:
            continue
        nmin, nmax = min(no, nr), max(no, nr)
        zero = math.ceil((nmax + 1) / 2)
        orig = zero - math.ceil((nmin - 1) / 2)
        fina = zero + math.floor((nmin - 1) / 2)
        X = np.fft.fftshift(A, axes=m)
        idx = [slice(None)] * A.ndim
        idx[m] = slice(orig - 1, fina)
        if no < nr:
            shp = list(A.shape)
            shp[m] = nr
            Y = np.zeros(shp, A.dtype)
            Y[tuple(idx)] = X
        else:
            Y = X[tuple(idx)]
        A = np.fft.ifftshift(Y, axes=m)
    return A


def level_data(yc, samp: "Sampling", H, N, n):
    """Compact data of a pyramid level, following the reference literally:
    zero-filled unitary k-space (solveXT.m:342 convention) -> image space ->
    resampling(y, NYRes) in image space (downsampleOperators.m:48-51) ->
    unitary FFT along the phase encodes at the level size (solveXT.m:342) ->
    buildHarmonicSampling of the k-space-cropped mask (downsampleOperators.m:36).
    yc: [Ns, N3, C] compact samples of the full level.  Returns
    (y_level [Ns', n3, C], k1, k2, off)."""
    N1, N2, N3 = N
    C = yc.shape[-1]
    Y = np.zeros((N1, N2, N3, C), np.complex128)
    Y[samp.k1, samp.k2] = yc
    y_img = np.fft.ifft(np.fft.ifft(Y, axis=0), axis=1) * math.sqrt(N1 * N2)
    y_lev = resample_fourier(y_img, tuple(n) + (C,))
    Yl = np.fft.fft(np.fft.fft(y_lev, axis=0), axis=1) / math.sqrt(n[0] * n[1])
    Hl = resample_kspace(H, n[:2])
    S = int(H.max())
    k1, k2, off, _ = build_harmonic_sampling(Hl, S)
    ns = off[S]
    return Yl[k1[:ns], k2[:ns]], k1[:ns], k2[:ns], off[:S + 1]


def level_size(N, l):
    """Methods/downsampleOperators.m:34: ceil(N/2^l)."""
    return [int(math.ceil(v / 2 ** l)) for v in N]


# ----------------------------------------------------------------------------
# Robust per-shot weights (Reconstruction/solveXT.m:455-468)
# ----------------------------------------------------------------------------
def prctile(x, p):
    """MATLAB prctile of a vector: the i-th smallest of n values sits at
    percentile 100 (i - 0.5) / n; linear interpolation in between, the
    minimum / maximum outside [100 * 0.5 / n, 100 (n - 0.5) / n]."""
    xs = np.sort(np.asarray(x, np.float64).ravel())
    n = len(xs)
    out = []
    for q in np.atleast_1d(p):
        pos = q / 100.0 * n + 0.5  # 1-based fractional index
        if pos <= 1:
            out.append(xs[0])
        elif pos >= n:
            out.append(xs[-1])
        else:
            i = int(math.floor(pos))
            f = pos - i
            out.append(xs[i - 1] + f * (xs[i] - xs[i - 1]))
    return np.array(out)


def robust_shot_weights(per_sample, state_of_sample, nshots, finer=True, perc=(0.125, 0.25), ener=0.95):
    """solveXT.m:455-468: per shot the percentiles (15:5:85 % when finer,
    else 30:5:70 %) of the log residual energies of its samples; across
    shots a robust std from the perc percentiles scaled by the normal
    quantiles and a robust mean from their mid percentile; per shot the mean
    z-score, and We = min((1 - normcdf(z)) / ((1 - ener) / nshots), 1).
    A shot without samples has NaN percentiles, which prctile across shots
    skips; its weight is 1 (it weights no sample)."""
    from scipy.special import erfcinv, ndtr, ndtri
    pu = np.arange(15, 86, 5) if finer else np.arange(30, 71, 5)
    logr = np.log(np.asarray(per_sample, np.float64))
    st = np.asarray(state_of_sample)
    rows, live = [], []
    for s in range(nshots):
        v = logr[st == s]
        if len(v):
            rows.append(prctile(v, pu))
            live.append(s)
    WeO = np.array(rows)
    lo = np.array([prctile(WeO[:, j], [perc[0] * 100, perc[1] * 100]) for j in range(len(pu))])  # [nperc, 2]
    Westd = (lo[:, 1] - lo[:, 0]) / (ndtri(perc[1]) - ndtri(perc[0]))
    pm = float(np.mean(perc))
    Wmea = np.array([prctile(WeO[:, j], [pm * 100])[0] for j in range(len(pu))])
    Wmea = Wmea + Westd * math.sqrt(2.0) * erfcinv(2 * pm)
    z = ((WeO - Wmea[None, :]) / Westd[None, :]).mean(axis=1)
    We = np.ones(nshots)
    We[live] = np.minimum((1.0 - ndtr(z)) / ((1.0 - ener) / nshots), 1.0)
    return We


# ----------------------------------------------------------------------------
# Motion binning (Methods/compressMotion.m:27-80)
# ----------------------------------------------------------------------------
def haar_dwt(x):
    """One level of MATLAB dwt(x, 'haar') in the default 'sym' extension
    mode: x is extended by one mirrored sample at each end, filtered with
    Lo_D = [1 1]/sqrt2 and Hi_D = [-1 1]/sqrt2 ('valid' convolution, length
    n + 1) and the even-indexed outputs are kept: a_k = (x_{2k-1} + x_{2k})
    / sqrt2, d_k = (x_{2k-1} - x_{2k}) / sqrt2 (1-based), with x_{n+1} = x_n
    for odd n."""
    x = [float(v) for v in x]
    n = len(x)
    ext = [x[0]] + x + [x[-1]]
    r = 1.0 / math.sqrt(2.0)
    a, d = [], []
    for k in range(2, n + 2, 2):  # 1-based indices of the 'valid' output kept
        lo = r * ext[k - 1] + r * ext[k]       # y(k) f(2) + y(k+1) f(1), f = Lo_D
        hi = r * ext[k - 1] - r * ext[k]       # f = Hi_D = [-r, r]
        a.append(lo)
        d.append(hi)
    return np.array(a), np.array(d)


def haar_idwt(a, d, n):
    """MATLAB idwt(a, d, 'haar', n): zero-interleave (dyadup), full
    convolution with Lo_R = [1 1]/sqrt2 and Hi_R = [1 -1]/sqrt2, keep the
    central n samples (for 2m -> 2m - 1 the last one is dropped)."""
    r = 1.0 / math.sqrt(2.0)
    out = []
    for ak, dk in zip(a, d):
        out.append(r * ak + r * dk)
        out.append(r * ak - r * dk)
    extra = len(out) - n
    first = extra // 2
    return np.array(out[first:first + n])


def wavedec_haar(x, NL):
    """MATLAB [C, L] = wavedec(x, NL, 'haar'): C = [a_NL, d_NL, ..., d_1],
    L = [len(a_NL), len(d_NL), ..., len(d_1), len(x)]."""
    a = np.asarray(x, np.float64)
    ds, lens = [], [len(a)]
    for _ in range(NL):
        a, d = haar_dwt(a)
        ds.append(d)
        lens.append(len(d))
    C = np.concatenate([a] + ds[::-1])
    L = [len(a)] + lens[1:][::-1] + [lens[0]]
    return C, L


def waverec_haar(C, L):
    """MATLAB waverec(C, L, 'haar')."""
    NL = len(L) - 2
    a = np.asarray(C[:L[0]], np.float64)
    pos = L[0]
    for j in range(NL):
        d = C[pos:pos + L[1 + j]]
        pos += L[1 + j]
        a = haar_idwt(a, d, L[2 + j])
    return a


def compress_motion(T, off, res, traLimX=(0.05, 0.02)):
    """compressMotion.m:27-80 for states that all carry motion (no outer
    segments): each of the 6 parameter series over the S states is
    decomposed to NL = floor(log2 S) Haar levels, coefficients below
    traLimX(1) res (translations, voxels) or traLimX(2) res degrees in
    radians (rotations) are zeroed and the series rebuilt (:30-36); runs of
    consecutive states whose rebuilt parameters agree to 1e-6 with the run's
    first state form one bin (:40-52) with the mean rebuilt parameters (:57-61)
    and the run's samples (their lists concatenated in state order).
    Returns (Tb [B x 6], off_b [B + 1], bin_of_state [S])."""
    T = np.atleast_2d(np.asarray(T, np.float64))
    S = T.shape[0]
    off = np.asarray(off)
    if S <= 1:
        return T.copy(), off.copy(), np.zeros(S, np.int64)
    NL = int(math.floor(math.log2(S)))
    Test = np.empty_like(T)
    for m in range(6):
        th = traLimX[0] * res if m < 3 else traLimX[1] * res * math.pi / 180.0
        C, L = wavedec_haar(T[:, m], NL)
        C = np.where(np.abs(C) < th, 0.0, C)
        Test[:, m] = waverec_haar(C, L)
    starts = []
    n = 0
    while n < S:
        m = 1
        while n + m < S and not np.any(np.abs(Test[n + m] - Test[n]) > 1e-6):
            m += 1
        starts.append((n, m))
        n += m
    Tb = np.array([Test[a:a + m].mean(axis=0) for a, m in starts])
    off_b = np.array([off[a] for a, _ in starts] + [off[S]])
    bin_of_state = np.concatenate([np.full(m, b) for b, (_, m) in enumerate(starts)])
    return Tb, off_b, bin_of_state


# ----------------------------------------------------------------------------
# Pyramid driver (Reconstruction/solveXT.m:266-531, joint CG + LM per level)
# ----------------------------------------------------------------------------
def tukey_iso(N, gibbs):
    """Methods/buildFilter.m:69-92,122 'tukeyIso': on the centred grid
    k_d = (2 pi / N_d)(j - ceil((N_d + 1)/2)), |k| = sqrt(sum k_d^2), window 1
    below pi alpha (alpha = 1 - gibbs), 0.5 (1 + cos(pi (|k| - pi alpha) /
    ((1 - alpha) pi))) up to pi, 0 from pi on; ifftshifted to DC first."""
    grids = [(2 * np.pi / n) * (np.arange(1, n + 1) - math.ceil((n + 1) / 2)) for n in N]
    kk = np.sqrt(sum(np.abs(g.reshape([-1 if d == m else 1 for d in range(len(N))])) ** 2
                     for m, g in enumerate(grids)))
    x = np.ones(tuple(N))
    alpha = 1.0 - gibbs
    if gibbs != 0:
        fkk = 0.5 * (1 + np.cos(np.pi * ((kk - np.pi * alpha) / ((1 - alpha) * np.pi))))
        x = np.where(kk >= np.pi * alpha, fkk, x)
    x = np.where(kk >= np.pi, 0.0, x)
    return np.fft.ifftshift(x)


def cubic_bspline(N):
    """Methods/buildFilter.m:102-106 'CubicBSpline' (ifftshifted): per
    dimension 6 sinc(k/(2 pi))^4 / (4 + 2 cos k), k = (2 pi/N)(j - ceil((N+1)/2))."""
    x = np.ones(tuple(N))
    for m, n in enumerate(N):
        k = (2 * np.pi / n) * (np.arange(1, n + 1) - math.ceil((n + 1) / 2))
        f = 6 * np.sinc(k / (2 * np.pi)) ** 4 / (4 + 2 * np.cos(k))
        x = x * f.reshape([-1 if d == m else 1 for d in range(len(N))])
    return np.fft.ifftshift(x)


def filter_level_rows(yl, l1, l2, n, gibbs):
    """filtering(yRes, tukeyIso) of downsampleOperators.m:54-57 for compact
    rows (phase encodes in k-space, readout in image space): FFT along the
    readout, window at (k1, k2, k3), inverse FFT."""
    H = tukey_iso(n, gibbs) if gibbs <= 1 else cubic_bspline(n)
    Y = np.fft.fft(yl, axis=1)
    Y = Y * H[np.asarray(l1), np.asarray(l2), :][:, :, None]
    return np.fft.ifft(Y, axis=1)


def _level_kspace_rows(k1, k2, off, N12, n12):
    """Full-resolution sample rows that survive the k-space crop of a level
    and their level indices (resampling(A, NYRes(1:2), 1) followed by
    buildHarmonicSampling), by signed frequency."""
    def sf(k, n):
        return np.where(k < (n + 1) // 2, k, k - n)
    f1, f2 = sf(np.asarray(k1), N12[0]), sf(np.asarray(k2), N12[1])
    keep = ((f1 >= -(n12[0] // 2)) & (f1 <= (n12[0] - 1) // 2) &
            (f2 >= -(n12[1] // 2)) & (f2 <= (n12[1] - 1) // 2))
    rows = np.nonzero(keep)[0]
    state = np.searchsorted(np.asarray(off), np.arange(len(k1)), side="right") - 1
    offl = np.zeros(len(off), np.int64)
    offl[1:] = np.cumsum(np.bincount(state[rows], minlength=len(off) - 1))
    return rows, np.mod(f1[rows], n12[0]), np.mod(f2[rows], n12[1]), offl


def pyramid_levels(N, nlevels):
    """Level sizes ceil(N/2^f), f = nlevels-1 .. 0 (downsampleOperators.m:34
    with isotropic voxels, pyramidPlan.m:19-20), and resIso
    (pyramidPlan.m:25, 1 mm voxels)."""
    out = []
    for l in range(nlevels):
        f = nlevels - 1 - l
        out.append((level_size(N, f), math.sqrt(2.0) * 2.0 ** f))
    return out


def cg_iterations(li, it, sizes, nX_prev, nIt2):
    """nX of Reconstruction/solveXT.m:417-425 while motion is estimated:
    level li (0-based, l = li + 1), outer iteration it (1-based), sizes[l]
    the level grids (resPyr(l) == resPyr(l-1) exactly when the grid
    repeats), nX_prev the value of the previous outer iteration of the
    level, nIt2 = nIt(2)."""
    if li == 0 and it == 1:
        return nIt2
    if li == 0:
        return max(nX_prev - 1, 1)
    same = list(sizes[li]) == list(sizes[li - 1])
    if not same and it == 1:
        return max(nIt2 - 1, 1)
    if not same:
        return max(nX_prev - 1, 1)
    if it == 1:
        return 0
    if it == 2:
        return max(nIt2 - 2, 1)
    return max(nX_prev - (it - 2), 1)


def pyramid_solve(yc, Smaps, k1, k2, off, N, nlevels=3, n_outer=3, n_cg=4, winit=1e-3,
                  traLimXT=(0.05, 0.02), echoesMin=4, final_cg=0, T0=None, callback=None, robust=True,
                  info=None, bin_motion=True, traLimX=(0.05, 0.02), eivaS=None, n_within=0, traj=None,
                  gibbs=0.05, accel=0, NY=None, cg_schedule="nX", fat_shift_dir="H", redFOV=1.0 / 3.0,
                  tolerSolve=1e-3, fraction_order=0.0):
    """Joint reconstruction over a spatial pyramid (solveXT.m:266-531) with
    the loop semantics kept: per level the operators are downsampled
    (downsampleOperators.m:34-60), x is resampled from the previous level,
    w = winit and cT are reset (:313-326), states with fewer than echoesMin
    samples are frozen (:322-324), tL = traLimXT * resIso(l) (:370); each
    outer iteration resets motion convergence on the schedule of :376-381,
    runs n_cg CG iterations with 'RelativeResidual2Error' and tol 0 (:431),
    then one LMsolver call (:507); when every state has converged, a final
    CG of final_cg iterations with 'Energy' tol 1e-3 closes the level when
    an outer iteration is left for it (:425-431 with estT(l) = 0 and nIt(1),
    then the break of :523).  With robust (AlignedRec = 2), at the
    full-resolution level the per-sample residuals after each CG give the
    robust per-shot weights EH.We used by the later CG calls (:441-481);
    info (a dict, optional) receives the last shot weights as info['We'].
    With bin_motion, every CG and the energies for the weights use the
    motion binned by compress_motion (res = resIso of the finest level,
    solveXT.m:392) while LMsolver keeps the per-shot motion (:496).  eivaS = (final,
    motion, image-while-estimating) coil counts: CG uses the first eivaS[2]
    coils (its preconditioner too), the closing CG eivaS[0], LMsolver
    eivaS[1] (:401-407, :501-503).  n_within full-resolution levels follow
    the spatial ones (pyramidPlan.m:21-22); at them, once a level has
    converged its estT(l) = 0 iteration (the closing CG, if final_cg) ends
    with the 15..85 % robust weights, every state of a shot weighted below
    one is flagged (:470-478), and the next level splits the flagged states
    at their middle profile (:270-287, traj from trajectory_states) and
    estimates only the new states on its first iteration (on all of them
    before the last level, :383-385).  info['sweep_of_state'] receives the
    shot of every final state.  gibbs: every level's data filtered by the
    radial Tukey window (downsampleOperators.m:54-57).  accel: the last
    accel levels have estT = 0 from the start (pyramidPlan.m:23): the closing
    CG only, with the last weights when leff > 0 (solveXT.m:412-414), then
    the outlier flags.  NY: the acquired phase-encode grid when it is smaller
    than N (SENSE folding, solveXT.m:198-200, :362-364): k1/k2 index it, each
    level's data live on ceil(NY/2^l) (downsampleOperators.m:34) and its
    encoding folds the level image onto that grid.  cg_schedule: 'nX' runs
    the CG iterations of solveXT.m:417-425 with nIt(2) = n_cg (none on the
    first outer iteration of a repeated full-resolution level), 'fixed' runs
    n_cg every time.  fat_shift_dir, redFOV: the readout ROI E.nF of LMsolver
    and the readout weights EH.Mbe of the energies behind the robust weights
    (solveXT.m:345-346, :443).  fraction_order: parXT.fractionOrder; nonzero
    gives LMsolver the weights E.Fd = buildFilter([n(1:2) numel(E.nF)],
    'FractionalFiniteDiscreteIso', N ./ n, order) at the samples
    (solveXT.m:353-354, :506).
    Returns (x at full size, T)."""
    from oracle import trajectory_oracle as TO
    NSw = len(off) - 1
    T = np.zeros((NSw, 6)) if T0 is None else np.array(T0, np.float64, copy=True)
    x = None
    res_fine = pyramid_levels(N, nlevels)[-1][1]
    ev = [Smaps.shape[3]] * 3 if eivaS is None else list(eivaS)
    sweep_of_state = list(range(NSw))
    kc1, kc2, offc = np.asarray(k1), np.asarray(k2), np.asarray(off)
    ysrc = yc
    if traj is not None:
        stateSample = np.array(traj["sweep_of_profile"])
        Ti = np.asarray(traj["time_index"])
        Ti = Ti[..., None] if Ti.ndim == 2 else Ti  # [NY1, NY2, repeats]
        H0 = np.asarray(traj["H"]).reshape(Ti.shape)
        hk1, hk2, hoff, hidx = build_harmonic_sampling(H0, NSw)
        if not (np.array_equal(hk1[:hoff[-2]], k1) and np.array_equal(hk2[:hoff[-2]], k2)):
            raise ValueError("traj does not describe the sampling k1, k2, off")
        row_of = {int(i): r for r, i in enumerate(hidx[:hoff[-2]])}  # linear index over [NY1, NY2, repeats]
    subT = [1] * NSw
    subTT = [1] * NSw
    last_W = None
    info = {} if info is None else info
    levels = pyramid_levels(N, nlevels) + [(list(N), res_fine)] * int(n_within)
    Lidx = len(levels) - 1
    NYg = list(N[:2]) if NY is None else [int(v) for v in NY[:2]]
    for li, (n, resIso) in enumerate(levels):
        leff = li - (nlevels - 1)
        ny = level_size(NYg, max(0, nlevels - 1 - li))
        if any(v == 2 for v in subT):
            T = np.concatenate([np.repeat(T[i:i + 1], subT[i], axis=0) for i in range(len(subT))])
            subTT = [v for v in subT for _ in range(v)]
            sweep_of_state = [sw for sw, v in zip(sweep_of_state, subT) for _ in range(v)]
            stateSample = TO.subdivide_states(stateSample, subT)
            H = np.zeros_like(Ti)
            for i in range(Ti.shape[0]):
                for j in range(Ti.shape[1]):
                    for r in range(Ti.shape[2]):
                        if Ti[i, j, r]:
                            H[i, j, r] = stateSample[Ti[i, j, r] - 1]
            kc1, kc2, offh, hidx = build_harmonic_sampling(H, len(sweep_of_state))
            offc = offh[:-1]
            kc1, kc2 = kc1[:offc[-1]], kc2[:offc[-1]]
            ysrc = yc[[row_of[int(i)] for i in hidx[:offc[-1]]]]
        subT = [1] * len(sweep_of_state)
        S = len(sweep_of_state)
        rows, l1, l2, loff = _level_kspace_rows(kc1, kc2, offc, NYg, ny)
        scale = math.sqrt(ny[0] * ny[1] / (NYg[0] * NYg[1]))
        yl = resample_fourier(ysrc[rows] * scale, (len(rows), n[2], yc.shape[2]))
        if gibbs:
            yl = filter_level_rows(yl, l1, l2, (ny[0], ny[1], n[2]), gibbs)
        Sl = resample_fourier(Smaps, tuple(n) + (Smaps.shape[3],))
        samp = Sampling(l1, l2, loff, ny[0], ny[1])
        full_res = list(n) == list(N)
        sweep_of_sample = np.array(sweep_of_state)[samp.state_of_sample()]

        def binned(T, We):
            if not bin_motion:
                return Sampling(l1, l2, loff, ny[0], ny[1], We), T
            Tb, offb, _ = compress_motion(T, loff, res_fine, traLimX)
            return Sampling(l1, l2, offb, ny[0], ny[1], We), Tb

        nF = readout_roi(n[2], fat_shift_dir, redFOV)
        Fd = None
        if fraction_order:
            if list(ny) != list(n[:2]):
                raise ValueError("fractional finite differences need unfolded data")
            F3 = fractional_filter([n[0], n[1], nF[1] - nF[0]], [N[d] / n[d] for d in range(3)], fraction_order)
            Fd = fractional_samples(F3, samp)
        Mbe = np.zeros(n[2])
        Mbe[nF[0]:nF[1]] = 1.0

        def weights(c, finer):
            samp_b, Tb = binned(T, samp.We)
            ps = compute_energy(yl[..., :c], x, Sl[..., :c], Sampling(l1, l2, samp_b.off, ny[0], ny[1]), Tb, N,
                                per_sample=True, Mbe=Mbe)
            Wshot = robust_shot_weights(ps, sweep_of_sample, NSw, finer=finer)
            samp.We = Wshot[sweep_of_sample]
            if info is not None:
                info["We"] = Wshot
            return Wshot
        Se = 1.0 / (np.sum(np.abs(Sl[..., :ev[2]]) ** 2, axis=-1) + 1e-9)
        Sef = 1.0 / (np.sum(np.abs(Sl[..., :ev[0]]) ** 2, axis=-1) + 1e-9)
        x = np.zeros(n, np.complex128) if x is None else resample_fourier(x, n)
        w = np.full(S, winit)
        nsamp = np.diff(loff)
        cT0 = nsamp < echoesMin
        cT = cT0.copy()
        tL = (traLimXT[0] * resIso, traLimXT[1] * resIso)
        nReset, cont = 0, 0
        nX = n_cg
        if li >= len(levels) - int(accel):
            if robust and leff > 0 and last_W is not None:
                samp.We = last_W[sweep_of_sample]
            if final_cg:
                samp_b, Tb = binned(T, samp.We)
                c = ev[0]
                x, _, _ = cg_solver(yl[..., :c], Sl[..., :c], samp_b, Tb, N, x0=x, nIt=final_cg,
                                    tolType="Energy", tol=tolerSolve, Se=Sef)
            if robust and leff > 0:
                Wshot = weights(ev[0], True)
                subT = [2 if Wshot[sw] < 1 else 1 for sw in sweep_of_state]
                if all(v == 1 for v in subT):
                    Lidx = li
            if callback is not None:
                callback(li, 1, x, T)
            if li == len(levels) - 1:
                subT = [1] * len(sweep_of_state)
            if info is not None and "We" in info:
                last_W = info["We"]
            continue
        for it in range(1, n_outer + 1):
            if (nReset == 0 and cont == 0) or (nReset > 0 and cont % nReset == 0) or cT.all():
                nReset += 1
                cont = 0
                cT = cT0.copy()
            if (it == 1 or li < Lidx) and any(v == 2 for v in subTT):
                cT = cT | (np.array(subTT) == 1)
            cont += 1
            samp_b, Tb = binned(T, samp.We)
            c = ev[2]
            nX = cg_iterations(li, it, [lv[0] for lv in levels], nX, n_cg) if cg_schedule == "nX" else n_cg
            if nX:
                x, _, _ = cg_solver(yl[..., :c], Sl[..., :c], samp_b, Tb, N, x0=x, nIt=nX,
                                    tolType="RelativeResidual2Error", tol=0.0, Se=Se)
            if robust and full_res:
                weights(c, leff > 0)
            T, w, cT, _ = lm_solver(yl[..., :ev[1]], x, Sl[..., :ev[1]], samp, T, N, w, cT, winit, tL,
                                    nF=None if nF == (0, n[2]) else nF, Fd=Fd)
            if callback is not None:
                callback(li, it, x, T)
            if cT.all():
                if it < n_outer:
                    if final_cg:
                        samp_b, Tb = binned(T, samp.We)
                        c = ev[0]
                        x, _, _ = cg_solver(yl[..., :c], Sl[..., :c], samp_b, Tb, N, x0=x, nIt=final_cg,
                                            tolType="Energy", tol=tolerSolve, Se=Sef)
                    if robust and leff > 0:
                        Wshot = weights(ev[0], True)
                        subT = [2 if Wshot[sw] < 1 else 1 for sw in sweep_of_state]
                        if all(v == 1 for v in subT):
                            Lidx = li
                break
        if li == len(levels) - 1:
            subT = [1] * len(sweep_of_state)
        if info is not None and "We" in info:
            last_W = info["We"]
    if info is not None:
        info["sweep_of_state"] = np.array(sweep_of_state)
    return resample_fourier(x, N), T
