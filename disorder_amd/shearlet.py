"""Shearlet-regularised final reconstruction (Reconstruction/CSsolver.m).

The shearlet system is the one CSsolver builds, buildShearlet(N, 1, gpu, 1,
'kos') (Shearlets/buildShearlet.m): one scale, shear parameter K = 1,
Kovacevic's smooth orthogonal 'kos' directional filter, the 'maxflat2'
scaling filter, real shearlets, full = 0.  Its 50 three-dimensional bands are
products of two 2-D filters on coordinate planes (combineShearlets,
buildShearlet.m:94-106); this module builds the 11 plane filters of each
plane on the host (cone 1 and cone 2 with shearings -2..2 and the low-pass,
buildShearlet.m:66-78 with Shearlets/buildWedgeBandpassAndLowpassFilters.m)
and hands the planes and the band table to the device
(dis_shearlet_create), which evaluates every band from them inside its FFT
passes.

CSsolver (CSsolver.m:31-100): from the image x of the aligned
reconstruction, IRLS weights We_w = (|IFFT3(conj(S_w) FFT3 x)| + 1e-2)^(p-2)
with p = 1; lambda = (2/p) tr_fid / tr_reg from Hutchinson trace estimates
(Reconstruction/traceEstimation.m: blocks of 4 normalised real Gaussian
probes until the standard error is below 1 % of the mean); then nOu rounds
of 3 preconditioned CG iterations ('None') on E^H We E + (la p/2) R_sh with
A.Se = 1/(sum_c |S_c|^2 + la p/2), re-weighting and re-estimating lambda
between rounds.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from . import ops

# -------------------------------------------------------------------- filters
_MAXFLAT2 = np.array([.0104933261758410, -.0263483047033631, -.0517766952966370, .276348304703363,
                      .582566738241592, .276348304703363, -.0517766952966369, -.0263483047033631,
                      .0104933261758408])


def _kos():
    """Shearlets/buildDirectionalFilter.m 'kos' (a0 = a1 = -sqrt3, a2 = 2 + sqrt3)."""
    r3 = math.sqrt(3.0)
    a0, a1, a2 = -r3, -r3, 2.0 + r3
    h = np.array([[0.0, -a1, -a0 * a1, 0.0], [-a2, -a0 * a2, -a0, 1.0], [0.0, a0 * a1 * a2, -a1 * a2, 0.0]])
    return h * (math.sqrt(2.0) / h.sum())


def _centre(x, shape):
    """resampling(x, shape, 2): centred zero-pad / crop with the centre at
    ceil((n + 1)/2) of the larger size, per dimension."""
    out = x
    for ax, nr in enumerate(shape):
        no = out.shape[ax]
        if no == nr:
            continue
        small, big = min(no, nr), max(no, nr)
        start = math.ceil((big + 1) / 2) - math.ceil((small - 1) / 2) - 1
        if no < nr:
            pad = [(0, 0)] * out.ndim
            pad[ax] = (start, nr - no - start)
            out = np.pad(out, pad)
        else:
            out = np.take(out, np.arange(start, start + nr), axis=ax)
    return out


def _zero_insert(x, ax, f):
    """Methods/upsampling.m: f - 1 zeros between samples along ax."""
    shp = list(x.shape)
    shp[ax] = f * shp[ax] - (f - 1)
    out = np.zeros(shp, dtype=x.dtype)
    sl = [slice(None)] * x.ndim
    sl[ax] = slice(None, None, f)
    out[tuple(sl)] = x
    return out


def _f2(x):
    """fft2D: ifftshift then FFT on both axes (centred array -> DC first)."""
    return np.fft.fft2(np.fft.ifftshift(x))


def _if2(x):
    return np.fft.fftshift(np.fft.ifft2(x))


def _plane_parts(Na, Nb):
    """(WD [Na, Nb, 5] for shearings +2..-2, BP [Na, Nb], LP [Na, Nb]) of
    buildWedgeBandpassAndLowpassFilters([Na Nb], K = 1)."""
    from scipy.signal import convolve2d
    sf = _MAXFLAT2 / _MAXFLAT2.sum()
    L = len(sf)
    wf = sf * (-1.0) ** (np.arange(L) - L // 2)        # mirrorModulation
    df = _kos()
    df = df / np.abs(df).sum()
    lp2 = np.convolve(sf, _zero_insert(sf, 0, 2))       # FLP2{1}
    BP = _f2(_centre(wf[None, :], (Na, Nb)))
    up = _zero_insert(_centre(convolve2d(_zero_insert(df, 0, 4), lp2[:, None]), (Na, Nb)), 1, 2)
    aux = _centre(sf[None, :], up.shape)
    up = _if2(_f2(up) * _f2(aux))
    aux_fl_f = _f2(aux[:, ::-1])
    rows = np.arange(Na) - Na // 2                       # r = i - floor(Na/2) - 1, 1-based i
    kk = np.fft.ifftshift((2 * np.pi / up.shape[1]) * (np.arange(1, up.shape[1] + 1) - math.ceil((up.shape[1] + 1) / 2)))
    upF = np.fft.fft(up, axis=1)
    WD = np.zeros((Na, Nb, 5), np.complex128)
    for l in range(-2, 3):
        shifted = np.fft.ifft(upF * np.exp(1j * l * rows[:, None] * kk[None, :]), axis=1)  # shearing(., l, [2 1])
        w = _if2(_f2(shifted) * aux_fl_f)
        WD[:, :, 2 - l] = _f2(2 * w[:, 0:2 * Nb:2])
    LP = _f2(_centre(np.outer(sf, sf), (Na, Nb)))
    return WD, BP, LP


def plane_filters(Na, Nb):
    """The 11 plane filters [Na, Nb, 11] (complex, DC first): cone 1 with
    shearings -2..2, cone 2 with shearings -2..2 (built on the transposed
    plane), then the low-pass (getShearlets with the full 2-D index list)."""
    lim = 9 + 2 * 8 + 2 * 4  # checkFilterFeasibility.m configuration 1 for K = 1
    if min(Na, Nb) < lim:
        raise ValueError(f"plane {Na}x{Nb} too small for the 'kos'/'maxflat2' shearlets (min {lim})")
    WD1, BP1, LP1 = _plane_parts(Na, Nb)
    WD2, BP2, _ = _plane_parts(Nb, Na)
    out = np.empty((Na, Nb, 11), np.complex128)
    for i, sh in enumerate(range(-2, 3)):
        out[:, :, i] = WD1[:, :, 2 - sh] * np.conj(BP1)
        out[:, :, 5 + i] = (WD2[:, :, 2 + sh] * np.conj(BP2)).T
    out[:, :, 10] = LP1
    return out


def band_table():
    """[50 x 4] (plane a, filter a, plane b, filter b) of the 3-D bands in
    getShearletIdxs(3, 1, full=0) order (planes 0 = (k1,k2), 1 = (k1,k3),
    2 = (k3,k2); filter index 5 (cone-1) + 2 + shearing, low-pass 10)."""
    rows = []
    for pyr in (1, 2, 3):
        for s1 in range(-2, 3):
            for s2 in range(-2, 3):
                keep = {1: True, 2: abs(s2) < 2, 3: abs(s1) < 2 and abs(s2) < 2}[pyr]
                if not keep:
                    continue
                if pyr == 1:
                    rows.append((0, 2 + s1, 2, 2 + s2))
                elif pyr == 2:
                    rows.append((1, 2 + s1, 2, 7 + s2))
                else:
                    rows.append((0, 7 + s1, 1, 7 + s2))
    rows.append((0, 10, 2, 10))
    return np.array(rows, np.int32)


# ------------------------------------------------------------------ device side
class ShearletSystem:
    """Device handle of the shearlet system of an n1 x n2 x n3 volume."""

    def __init__(self, n, ctx: ops.Context = None):
        self.ctx = ctx or ops.default_context()
        self.n = tuple(int(v) for v in n)
        n1, n2, n3 = self.n
        planes = [plane_filters(n1, n2), plane_filters(n1, n3), plane_filters(n3, n2)]
        self._planes = [np.asfortranarray(p.astype(np.complex64)) for p in planes]
        self.bands = band_table()
        h = ctypes.c_void_p()
        a = (ctypes.c_int32 * 3)(*self.n)
        bt = np.ascontiguousarray(self.bands, np.int32)
        self.ctx._prep()
        ops.check(self.ctx.lib.dis_shearlet_create(self.ctx.h, a, *[p.ctypes.data_as(ctypes.c_void_p) for p in
                                                                   self._planes],
                                                   11, bt.ctypes.data_as(_lib.c_int32_p), len(bt), ctypes.byref(h)),
                  self.ctx.h)
        self.h = h
        self.nw = len(bt)

    def close(self):
        if self.h:
            self.ctx.lib.dis_shearlet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self, x, epsil=1e-2, p=1.0):
        self.ctx._prep()
        ops.check(self.ctx.lib.dis_shearlet_weights(self.ctx.h, self.h, ops._ptr(x), float(epsil), float(p)),
                  self.ctx.h)

    def regularize(self, x, scale, out=None):
        """out + scale * R_sh(x) (a new zero volume when out is None)."""
        if out is None:
            out = ops.empty_colmajor(self.n, x.device)
            out.zero_()
        self.ctx._prep()
        ops.check(self.ctx.lib.dis_shearlet_regularize(self.ctx.h, self.h, ops._ptr(x), float(scale), ops._ptr(out)),
                  self.ctx.h)
        return out

    def energy(self, x):
        r = ctypes.c_double(0.0)
        self.ctx._prep()
        ops.check(self.ctx.lib.dis_shearlet_energy(self.ctx.h, self.h, ops._ptr(x), ctypes.byref(r)), self.ctx.h)
        return r.value

    def fidelity(self, S, Wgrid, x):
        r = ctypes.c_double(0.0)
        self.ctx._prep()
        ops.check(self.ctx.lib.dis_fidelity_trace(self.ctx.h, self.h, ops._ptr(S), int(S.shape[3]),
                                                  ops._fptr(Wgrid), ops._ptr(x), ctypes.byref(r)), self.ctx.h)
        return r.value


def fidelity_folded(S, Wgrid, x, NY, ctx):
    """One fidelity trace sample with SENSE folding onto the acquired grid NY
    (traceEstimation.m:61-79 with E.Uf): dis_fidelity_trace_folded."""
    n = (ctypes.c_int32 * 3)(*[int(v) for v in x.shape[:3]])
    r = ctypes.c_double(0.0)
    ctx._prep()
    ops.check(ctx.lib.dis_fidelity_trace_folded(ctx.h, n, ops._ptr(S), int(S.shape[3]), int(NY[0]), int(NY[1]),
                                                ops._fptr(Wgrid), ops._ptr(x), ctypes.byref(r)), ctx.h)
    return r.value


def grid_weights(k1, k2, We, n12):
    """EH.We on the k-space grid (traceEstimation.m:43-48): the weight of the
    sample at each sampled (k1, k2), 1 elsewhere."""
    W = np.ones(n12, np.float32)
    if We is not None:
        W[np.asarray(k1), np.asarray(k2)] = np.asarray(We, np.float32)
    return W


def trace_estimation(sample, shape, rng, device, tol=1e-2, block=4, max_probes=4096):
    """traceEstimation.m:52-110: mean of sample(probe) over blocks of
    `block` real N(0,1) probes normalised to unit mean square, until
    std/sqrt(n) < tol * mean (at least two probes).  Returns (mean, n).
    A probe value that is not finite, or max_probes probes without
    reaching the tolerance, raises (the reference's loop would not end)."""
    import torch
    vals = []
    while True:
        if len(vals) >= max_probes:
            raise RuntimeError(f"trace estimation did not reach tol {tol} in {max_probes} probes")
        for _ in range(block):
            z = rng.standard_normal(shape)
            z = z / math.sqrt(float(np.sum(z * z)) / z.size)
            zt = ops.colmajor(torch.from_numpy(np.asfortranarray(z.astype(np.complex64))).to(device))
            v = float(sample(zt))
            if not math.isfinite(v):
                raise RuntimeError("trace estimation: probe value is not finite")
            vals.append(v)
        n = len(vals)
        m = float(np.mean(vals))
        if n >= 2 and float(np.std(vals, ddof=1)) / math.sqrt(n) / (m + np.finfo(np.float64).eps) < tol:
            return m, n


def CSsolver(y, E: ops.Encoding, T, x, S, k1, k2, We=None, nOu=2, p=1.0, epsil=1e-2, tol=1e-2, seed=0,
             system: ShearletSystem = None, info=None):
    """Reconstruction/CSsolver.m:31-100 from the image x (x given, so no
    initial CG).  y, S, x on the device (column-major complex64); E the
    encoding with the sample weights We (per sample, or None); k1, k2 the
    sample positions for the weights on the grid.  Returns (x, lambdas)."""
    import torch
    rng = np.random.default_rng(seed)
    n = tuple(int(v) for v in x.shape)
    sh = system or ShearletSystem(n, E.ctx)
    NY = tuple(getattr(E, "NY", n[:2]))  # acquired phase-encode grid (SENSE folding below n[:2])
    Wgrid = torch.from_numpy(grid_weights(k1, k2, We, NY).reshape(-1, order="F")).to(x.device)
    x = ops.colmajor(x.clone())
    sh.weights(x, epsil, p)
    if NY != n[:2]:
        tr_fid, nf = trace_estimation(lambda z: fidelity_folded(S, Wgrid, z, NY, E.ctx), n, rng, x.device, tol)
    else:
        tr_fid, nf = trace_estimation(lambda z: sh.fidelity(S, Wgrid, z), n, rng, x.device, tol)
    tr_reg, nr = trace_estimation(sh.energy, n, rng, x.device, tol)
    la = (2.0 / p) * tr_fid / (tr_reg + np.finfo(np.float64).eps)
    lambdas = [la]
    Tf, Tp = E._T(T)
    for it in range(1, nOu + 1):
        Se = ops.coil_precond(S, la * p / 2, E.ctx)
        nn = ctypes.c_int32(0)
        E.ctx._prep()
        ops.check(E.ctx.lib.dis_cg_cs(E.h, Tp, ops._ptr(y), ops._fptr(Se), None, 0.0, sh.h, float(la * p / 2),
                                      ops._ptr(x), 1, 3, ops.TOLTYPES["None"], float(tol), ctypes.byref(nn)), E.ctx.h)
        if it != nOu:
            sh.weights(x, epsil, p)
            tr_reg, _ = trace_estimation(sh.energy, n, rng, x.device, tol)
            la = (2.0 / p) * tr_fid / (tr_reg + np.finfo(np.float64).eps)
            lambdas.append(la)
    if info is not None:
        info.update({"probes_fid": nf, "probes_reg": nr})
    if system is None:
        sh.close()
    return x, lambdas
