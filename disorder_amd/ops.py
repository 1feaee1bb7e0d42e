"""Host-side mirror of DISORDER's hot-path interface over the C-ABI.

Arrays use MATLAB's memory layout: a volume is a complex64 torch tensor of
shape (n1, n2, n3) with column-major strides (see ``colmajor``); per-state
stacks are (n1, n2, n3, S); compact k-space data is (Ns, n3, C).  Device
tensors are passed to the library as raw pointers; all work runs in
libdisorder_hip.so on the context's stream.

Function names follow the reference: ``sincRigidTransform``
(Methods/sincRigidTransform.m:1), ``encode`` (Reconstruction/encode.m:1),
``decode`` (Reconstruction/decode.m:1), ``encodedecode``
(Reconstruction/encodedecode.m:1), ``computeEnergy``
(Reconstruction/computeEnergy.m:1).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import DisGeom, check


def colmajor(t: torch.Tensor) -> torch.Tensor:
    """Same values and shape, column-major (MATLAB) strides."""
    nd = t.dim()
    perm = list(range(nd))[::-1]
    return t.permute(*perm).contiguous().permute(*perm)


def is_colmajor(t: torch.Tensor) -> bool:
    st = 1
    for d in range(t.dim()):
        if t.shape[d] != 1 and t.stride(d) != st:
            return False
        st *= t.shape[d]
    return True


def empty_colmajor(shape, device, dtype=torch.complex64):
    return torch.empty(tuple(shape)[::-1], dtype=dtype, device=device).permute(*range(len(shape) - 1, -1, -1))


def _ptr(t: torch.Tensor):
    if t.dtype != torch.complex64:
        raise TypeError("complex64 tensors are required")
    if not t.is_cuda:
        raise TypeError("device tensors are required")
    if not is_colmajor(t):
        raise ValueError("tensors must have column-major (MATLAB) strides; use colmajor()")
    return ctypes.c_void_p(t.data_ptr())


def _fptr(t):
    """Real float32 column-major device tensor (or None) -> float*."""
    if t is None:
        return None
    if t.dtype != torch.float32 or not t.is_cuda or not is_colmajor(t):
        raise TypeError("float32 column-major device tensors are required")
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), _lib.c_float_p)


def _tbuf(T, rows=None):
    """Motion [S x 6] -> column-major float64 host buffer."""
    T = np.atleast_2d(np.asarray(T, dtype=np.float64))
    if T.shape[1] != 6:
        raise ValueError("T must be [S x 6]")
    flat = np.ascontiguousarray(T.T).ravel()  # T[s + S*g]
    return flat, flat.ctypes.data_as(_lib.c_double_p)


class Context:
    """One GPU, one stream (dis_ctx)."""

    def __init__(self, device: int = 0, workspace_limit: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.dis_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = torch.device("cuda", device)
        if workspace_limit:
            check(self.lib.dis_set_workspace_limit(self.h, workspace_limit), self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.dis_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.lib.dis_sync(self.h), self.h)

    def set_cache_budget(self, nbytes):
        """Bytes of state volumes a rigid transform keeps in flight
        (dis_set_cache_budget; 0 runs every state in every pass launch)."""
        check(self.lib.dis_set_cache_budget(self.h, ctypes.c_int64(int(nbytes))), self.h)

    def profile(self, on=True):
        check(self.lib.dis_profile_enable(self.h, 1 if on else 0), self.h)

    def profile_read(self):
        """{kernel class: (launches, total ms, algorithmic bytes)}"""
        out = {}
        for k, name in _lib.K_NAMES.items():
            n = ctypes.c_int64(0)
            ms = ctypes.c_double(0)
            by = ctypes.c_double(0)
            check(self.lib.dis_profile_read(self.h, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)), self.h)
            out[name] = (n.value, ms.value, by.value)
        return out

    def _prep(self):
        # order the library's stream after torch's producers of the inputs
        torch.cuda.current_stream(self.device).synchronize()


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def sincRigidTransform(x: torch.Tensor, T, di: int = 1, sumR: bool = True, Nfull=None, ctx: Context = None):
    """Methods/sincRigidTransform.m:1 with the factors of
    precomputeFactorsSincRigidTransform(kG, rkG, T, di, 0, 1, 1).

    di=1: x (n1,n2,n3) or (n1,n2,n3,S) -> (n1,n2,n3,S).
    di=0: x (n1,n2,n3,S) -> (n1,n2,n3) if sumR else (n1,n2,n3,S)."""
    ctx = ctx or default_context(x.device.index or 0)
    Tf, Tp = _tbuf(T)
    S = Tf.size // 6
    n = (ctypes.c_int32 * 3)(*x.shape[:3])
    N = (ctypes.c_int32 * 3)(*(Nfull if Nfull is not None else x.shape[:3]))
    if di == 1:
        per_state = 1 if x.dim() == 4 and x.shape[3] > 1 else 0
        if x.dim() == 4 and x.shape[3] not in (1, S):
            raise ValueError("state count of x does not match T")
        out = empty_colmajor(tuple(x.shape[:3]) + (S,), x.device)
    else:
        if x.dim() != 4 or x.shape[3] != S:
            raise ValueError("x must be (n1,n2,n3,S) for the inverse")
        per_state = 1
        out = empty_colmajor(tuple(x.shape[:3]) + ((S,) if not sumR else ()), x.device)
    ctx._prep()
    check(ctx.lib.dis_rigid_transform(ctx.h, n, N, Tp, S, di, int(bool(sumR)), per_state, _ptr(x), _ptr(out)),
          ctx.h)
    return out


def sincRigidTransformGradient(x: torch.Tensor, T, Nfull=None, ctx: Context = None):
    """Methods/sincRigidTransformGradient.m:1 with the factors of
    precomputeFactorsSincRigidTransform(kG, rkG, T, 1, 1, 1, 1): returns
    (T x as (n1,n2,n3,S), dT x/dT_g as (n1,n2,n3,S,6))."""
    ctx = ctx or default_context(x.device.index or 0)
    Tf, Tp = _tbuf(T)
    S = Tf.size // 6
    n = (ctypes.c_int32 * 3)(*x.shape[:3])
    N = (ctypes.c_int32 * 3)(*(Nfull if Nfull is not None else x.shape[:3]))
    Tx = empty_colmajor(tuple(x.shape[:3]) + (S,), x.device)
    G = empty_colmajor(tuple(x.shape[:3]) + (S, 6), x.device)
    ctx._prep()
    check(ctx.lib.dis_rigid_transform_grad(ctx.h, n, N, Tp, S, _ptr(x), _ptr(Tx), _ptr(G)), ctx.h)
    return Tx, G


class Encoding:
    """The E / EH structures of the reconstruction path as one plan
    (dis_plan): coil maps Sf, per-segment sampling (E.Fs rows as FFT-order
    indices k1/k2, offsets E.nSt), weights EH.We and NMs motion states;
    NY (the acquired phase-encode grid, E.Uf{1:2}.NY) below the maps' grid
    turns on SENSE folding, with k1/k2 indexing the NY grid."""

    def __init__(self, Smaps: torch.Tensor, k1, k2, state_off, Nfull=None, We=None, nmotion=None,
                 ctx: Context = None, NY=None):
        self.ctx = ctx or default_context(Smaps.device.index or 0)
        self.S = Smaps  # kept alive: the plan borrows the pointer
        n = tuple(Smaps.shape[:3])
        C = Smaps.shape[3]
        self.n = n
        self.Nfull = tuple(int(v) for v in Nfull) if Nfull is not None else n
        self.C = C
        self.k1 = np.ascontiguousarray(k1, dtype=np.int32)
        self.k2 = np.ascontiguousarray(k2, dtype=np.int32)
        self.off = np.ascontiguousarray(state_off, dtype=np.int32)
        nst = len(self.off) - 1
        self.nstates = nst
        self.nmotion = nst if nmotion is None else int(nmotion)
        self.Ns = int(self.off[-1])
        g = DisGeom()
        g.n[:] = n
        g.Nfull[:] = tuple(Nfull) if Nfull is not None else n
        g.ncoils = C
        g.nstates = nst
        g.nmotion = self.nmotion
        g.nsamples = self.Ns
        self.geom = g
        wp = None
        if We is not None:
            self.We = np.ascontiguousarray(We, dtype=np.float32)
            wp = self.We.ctypes.data_as(_lib.c_float_p)
        h = ctypes.c_void_p()
        check(self.ctx.lib.dis_plan_create(self.ctx.h, ctypes.byref(g), _ptr(Smaps),
                                           self.k1.ctypes.data_as(_lib.c_int32_p),
                                           self.k2.ctypes.data_as(_lib.c_int32_p),
                                           self.off.ctypes.data_as(_lib.c_int32_p), wp, ctypes.byref(h)),
              self.ctx.h)
        self.h = h
        # SENSE folding (E.Uf / EH.Ub): data acquired on an NY[0] x NY[1]
        # phase-encode grid smaller than the reconstruction grid
        self.NY = tuple(int(v) for v in NY[:2]) if NY is not None else n[:2]
        if self.NY != n[:2]:
            check(self.ctx.lib.dis_plan_set_fold(self.h, self.NY[0], self.NY[1]), self.ctx.h)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.dis_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, We):
        if We is None:
            check(self.ctx.lib.dis_plan_set_weights(self.h, None), self.ctx.h)
        else:
            self.We = np.ascontiguousarray(We, dtype=np.float32)
            check(self.ctx.lib.dis_plan_set_weights(self.h, self.We.ctypes.data_as(_lib.c_float_p)), self.ctx.h)

    def set_readout_roi(self, nF):
        """E.nF (LMsolver.m:33,68,98): the readout planes [lo, hi) (0-based)
        over which LMsolver and lm_partials sum; None for every plane."""
        lo, hi = (0, 0) if nF is None else (int(nF[0]), int(nF[1]))
        check(self.ctx.lib.dis_plan_set_readout_roi(self.h, lo, hi), self.ctx.h)
        self.nF = None if nF is None else (lo, hi)

    def set_lm_filter(self, Fd):
        """E.Fd (Reconstruction/LMsolver.m:88-91,101-104,186-189): per-sample
        fractional finite-difference weights [Ns x nF] along the readout DFT
        of the ROI (nF = its length, n3 without one), applied by LMsolver
        and lm_partials; None clears."""
        if Fd is None:
            check(self.ctx.lib.dis_plan_set_lm_filter(self.h, None, 0), self.ctx.h)
            self.Fd = None
            return
        Fd = np.asfortranarray(np.asarray(Fd, dtype=np.float32))
        if Fd.ndim != 2 or Fd.shape[0] != self.Ns:
            raise ValueError("Fd must be [nsamples x nF]")
        self.Fd = Fd
        check(self.ctx.lib.dis_plan_set_lm_filter(self.h, Fd.ctypes.data_as(_lib.c_float_p), Fd.shape[1]),
              self.ctx.h)

    def _T(self, T):
        if self.nmotion == 0:
            return np.zeros(6), None
        Tf, Tp = _tbuf(T)
        if Tf.size != 6 * self.nmotion:
            raise ValueError("T must be [nmotion x 6]")
        return Tf, Tp

    def encode(self, x, T, out=None):
        Tf, Tp = self._T(T)
        y = out if out is not None else empty_colmajor((self.Ns, self.n[2], self.C), x.device)
        self.ctx._prep()
        check(self.ctx.lib.dis_encode(self.h, Tp, _ptr(x), _ptr(y)), self.ctx.h)
        return y

    def decode(self, y, T, out=None):
        Tf, Tp = self._T(T)
        x = out if out is not None else empty_colmajor(self.n, y.device)
        self.ctx._prep()
        check(self.ctx.lib.dis_decode(self.h, Tp, _ptr(y), _ptr(x)), self.ctx.h)
        return x

    def encodedecode(self, x, T, out=None):
        Tf, Tp = self._T(T)
        o = out if out is not None else empty_colmajor(self.n, x.device)
        self.ctx._prep()
        check(self.ctx.lib.dis_encodedecode(self.h, Tp, _ptr(x), _ptr(o)), self.ctx.h)
        return o

    def energy(self, x, y, T, Mbe=None, per_sample=False):
        Tf, Tp = self._T(T)
        ps = np.zeros(max(1, self.Ns), np.float64)
        tot = ctypes.c_double(0.0)
        mb = None
        if Mbe is not None:
            self._Mbe = np.ascontiguousarray(Mbe, dtype=np.float32)
            mb = self._Mbe.ctypes.data_as(_lib.c_float_p)
        self.ctx._prep()
        check(self.ctx.lib.dis_energy(self.h, Tp, _ptr(x), _ptr(y), mb, ps.ctypes.data_as(_lib.c_double_p),
                                      ctypes.byref(tot)), self.ctx.h)
        return ps[:self.Ns] if per_sample else tot.value


TOLTYPES = {"RelativeResidual2Error": 1, "Energy": 6, "None": 7, "MaxUpdate": 8}


def CGsolver(y, E: "Encoding", T, x=None, nIt=300, tolType="Energy", tol=1e-5, Se=None, Ma=None, lam=0.0):
    """Reconstruction/CGsolver.m:1 -- preconditioned CG on E^H We E x + lam x =
    E^H We y with preconditioner Se (A.Se) and mask Ma (C.Ma); FP64 scalars.
    Returns (x, En, nn): En is (2, k) [fidelity; regulariser] of the odd
    iterations for tolType 'Energy'."""
    tt = TOLTYPES[tolType]
    init = x is not None
    xo = colmajor(x.clone()) if init else empty_colmajor(E.n, y.device)
    En = np.zeros(2 * (nIt // 2 + 2), np.float64)
    nEn = ctypes.c_int32(0)
    nn = ctypes.c_int32(0)
    Tf, Tp = E._T(T)
    E.ctx._prep()
    check(E.ctx.lib.dis_cg(E.h, Tp, _ptr(y), _fptr(Se), _fptr(Ma), float(lam), _ptr(xo), 1 if init else 0, int(nIt),
                           tt, float(tol), En.ctypes.data_as(_lib.c_double_p), ctypes.byref(nEn), ctypes.byref(nn)),
          E.ctx.h)
    return xo, En[:2 * nEn.value].reshape(-1, 2).T.copy(), nn.value


def autotune(n, Nfull=None, S=16, ctx: Context = None):
    """Pick the pencil-pass shape of every axis of grid n by timing the
    candidates (dis_autotune).  Returns [{generic, P, radix, ms}] per axis."""
    ctx = ctx or default_context()
    a = (ctypes.c_int32 * 3)(*[int(v) for v in n])
    b = (ctypes.c_int32 * 3)(*[int(v) for v in (Nfull or n)])
    ch = (ctypes.c_int32 * 9)()
    ms = (ctypes.c_double * 3)()
    check(ctx.lib.dis_autotune(ctx.h, a, b, int(S), ch, ms), ctx.h)
    return [{"generic": bool(ch[3 * d]), "P": ch[3 * d + 1], "radix": ch[3 * d + 2], "ms": ms[d]} for d in range(3)]


def CGsolver_batched(ys, Es, Ts, Ses=None, nIt=12, use_graph=True, ctx: Context = None):
    """Batched CG over independent reconstructions (dis_cg_batched): volume i
    solves E_i^H E_i x = E_i^H y_i with preconditioner Ses[i], x0 = 0, exactly
    nIt iterations ('RelativeResidual2Error', tol 0).  Returns (xs,
    res [nvol x nIt], ms [setup, first iteration, remaining iterations])."""
    ctx = ctx or Es[0].ctx
    nv = len(Es)
    xs = [empty_colmajor(E.n, ys[i].device) for i, E in enumerate(Es)]
    Tb = [E._T(T) for E, T in zip(Es, Ts)]
    plans = (ctypes.c_void_p * nv)(*[E.h.value for E in Es])
    Tp = (_lib.c_double_p * nv)(*[tp for _, tp in Tb])
    yp = (ctypes.c_void_p * nv)(*[_ptr(y).value for y in ys])
    xp = (ctypes.c_void_p * nv)(*[_ptr(x).value for x in xs])
    sp = None
    if Ses is not None:
        sp = (_lib.c_float_p * nv)(*[_fptr(s) if s is not None else None for s in Ses])
    res = np.zeros((nv, max(1, nIt)), np.float64)
    ms = np.zeros(3, np.float64)
    ctx._prep()
    check(ctx.lib.dis_cg_batched(ctx.h, nv, plans, Tp, yp, sp, xp, int(nIt), 1 if use_graph else 0,
                                 res.ctypes.data_as(_lib.c_double_p), ms.ctypes.data_as(_lib.c_double_p)), ctx.h)
    return xs, res[:, :nIt], ms


def lm_partials(y, E: "Encoding", x, T):
    """Reconstruction/LMsolver.m:52-122: per motion state the residual energy
    EnPrev (divided by numel(y)), gradient dG (S,6) and Hessian entries dH
    (S,21, pairs ha of LMsolver.m:26)."""
    Tf, Tp = E._T(T)
    sums = np.zeros((E.nmotion, 28), np.float64)
    E.ctx._prep()
    check(E.ctx.lib.dis_lm_partials(E.h, Tp, _ptr(x), _ptr(y), sums.ctypes.data_as(_lib.c_double_p)), E.ctx.h)
    nel = E.Ns * E.n[2] * E.C
    return sums[:, 0] / nel, sums[:, 1:7].copy(), sums[:, 7:].copy()


def LMsolver(y, E: "Encoding", x, T, w, cT, winit=1e-3, tL=(0.05, 0.02)):
    """Reconstruction/LMsolver.m:1 -- one Levenberg-Marquardt update of the
    motion states.  Returns (T, w, cT, En, n_trials)."""
    Tf, Tp = E._T(T)
    wv = np.ascontiguousarray(w, dtype=np.float64).copy()
    cv = np.ascontiguousarray(cT, dtype=np.uint8).copy()
    tl = np.ascontiguousarray(tL, dtype=np.float64)
    En = np.zeros(E.nmotion, np.float64)
    nt = ctypes.c_int32(0)
    E.ctx._prep()
    check(E.ctx.lib.dis_lm(E.h, Tp, wv.ctypes.data_as(_lib.c_double_p), cv.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                           float(winit), tl.ctypes.data_as(_lib.c_double_p), _ptr(x), _ptr(y),
                           En.ctypes.data_as(_lib.c_double_p), ctypes.byref(nt)), E.ctx.h)
    Tn = Tf.reshape(6, E.nmotion).T.copy()
    return Tn, wv, cv.astype(bool), En, nt.value


def encode(x, E: Encoding, T):
    """Reconstruction/encode.m:1 -- y = E x."""
    return E.encode(x, T)


def decode(y, E: Encoding, T):
    """Reconstruction/decode.m:1 -- x = E^H (We .* y)."""
    return E.decode(y, T)


def encodedecode(x, E: Encoding, T):
    """Reconstruction/encodedecode.m:1 -- E^H We E x."""
    return E.encodedecode(x, T)


def computeEnergy(y, x, E: Encoding, T, Mbe=None, per_sample=False):
    """Reconstruction/computeEnergy.m:34-57 (no=2, FP64 accumulation)."""
    return E.energy(x, y, T, Mbe=Mbe, per_sample=per_sample)


# ---------------------------------------------------------------- coil compression
def _coil_view(t, name):
    """[..., C] column-major complex64 device tensor -> (nv, C)."""
    if t.dim() < 2:
        raise ValueError(f"{name} must have a trailing coil dimension")
    _ptr(t)
    C = t.shape[-1]
    return t.numel() // max(C, 1), C


def coil_gram(S, ctx: Context = None):
    """Methods/compressCoils.m:78-87: P = sum_vox Shat S^H with
    Shat = S / (sum_c |S_c|^2 + 1e-9); complex128 numpy [C x C]."""
    ctx = ctx or default_context(S.device.index or 0)
    nv, C = _coil_view(S, "S")
    P = np.zeros((C, C), np.complex128, order="F")
    ctx._prep()
    check(ctx.lib.dis_coil_gram(ctx.h, _ptr(S), nv, C, P.ctypes.data_as(_lib.c_double_p)), ctx.h)
    return P


def coil_basis(P, perc):
    """Methods/compressCoils.m:93-111: (U, D, C) from svd(P) on the host in FP64."""
    P = np.asfortranarray(P, dtype=np.complex128)
    C = P.shape[0]
    pv = np.atleast_1d(np.asarray(perc, np.float64))
    cc = np.zeros(len(pv), np.int32)
    D = np.zeros(C, np.float64)
    U = np.zeros((C, C), np.complex128, order="F")
    lib = _lib.load()
    check(lib.dis_coil_basis(P.ctypes.data_as(_lib.c_double_p), C, pv.ctypes.data_as(_lib.c_double_p), len(pv),
                             cc.ctypes.data_as(_lib.c_int32_p), D.ctypes.data_as(_lib.c_double_p),
                             U.ctypes.data_as(_lib.c_double_p)))
    return U, D, cc


def coil_apply(A, x, terms=3, ctx: Context = None):
    """aplGPU(A, x, 4) of Methods/compressCoils.m:117-118: contracts the
    trailing coil dimension of x with A [Cc x C] on the matrix cores."""
    ctx = ctx or default_context(x.device.index or 0)
    nv, C = _coil_view(x, "x")
    A = np.asfortranarray(A, dtype=np.complex128)
    if A.shape[1] != C:
        raise ValueError("A must be [Cc x C]")
    Cc = A.shape[0]
    out = empty_colmajor(tuple(x.shape[:-1]) + (Cc,), x.device)
    ctx._prep()
    check(ctx.lib.dis_coil_apply(ctx.h, A.ctypes.data_as(_lib.c_double_p), Cc, C, _ptr(x), nv, _ptr(out),
                                 int(terms)), ctx.h)
    return out


def compressCoils(S, perc, y=None, terms=3, ctx: Context = None):
    """Methods/compressCoils.m:1 (useFold=0, useInvMot=0):
    [S, y, C, D, U] = compressCoils(S, perc, y).  S [n1,n2,n3,C] and y [...,C]
    are column-major complex64 device tensors; perc is an energy fraction
    (< 1) or a component count (>= 1), or a list of them.  The compression
    matrix is A = U(:, 1:max(C))^H.  ``terms`` picks the split-bf16 product
    count of the MFMA contraction (3 or 6)."""
    ctx = ctx or default_context(S.device.index or 0)
    if perc is None:
        return S, y, np.array([S.shape[-1]]), None, None
    P = coil_gram(S, ctx)
    U, D, cc = coil_basis(P, perc)
    Cc = int(cc.max())
    A = np.conj(U[:, :Cc]).T
    Sc = coil_apply(A, S, terms, ctx)
    yc = None if y is None else coil_apply(A, y, terms, ctx)
    return Sc, yc, cc, D, U


# ---------------------------------------------------------------- pyramid levels
def resampling(x, Nres, ctx: Context = None):
    """Methods/resampling.m:1 (fo=0, mirror=0) on the GPU: x [n1,n2,n3] or
    [n1,n2,n3,B] column-major complex64 -> [Nres..., B]."""
    ctx = ctx or default_context(x.device.index or 0)
    _ptr(x)
    nin = tuple(int(v) for v in x.shape[:3])
    nout = tuple(int(v) for v in Nres[:3])
    batch = int(np.prod(x.shape[3:])) if x.dim() > 3 else 1
    out = empty_colmajor(nout + tuple(x.shape[3:]), x.device)
    a = (ctypes.c_int32 * 3)(*nin)
    b = (ctypes.c_int32 * 3)(*nout)
    ctx._prep()
    check(ctx.lib.dis_resample(ctx.h, a, b, batch, _ptr(x), _ptr(out)), ctx.h)
    return out


def gather_samples(y, rows, scale=1.0, ctx: Context = None):
    """Rows of compact data y [Ns, n3, C] -> [len(rows), n3, C], times scale."""
    ctx = ctx or default_context(y.device.index or 0)
    _ptr(y)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    Ns = int(y.shape[0])
    ncol = int(np.prod(y.shape[1:]))
    out = empty_colmajor((len(rows),) + tuple(y.shape[1:]), y.device)
    ctx._prep()
    check(ctx.lib.dis_gather_samples(ctx.h, _ptr(y), Ns, ncol, rows.ctypes.data_as(_lib.c_int32_p), len(rows),
                                     float(scale), _ptr(out)), ctx.h)
    return out


def filter_rows(y, k1, k2, n, gibbs, ctx: Context = None):
    """Methods/downsampleOperators.m:54-57: the radial Tukey window
    buildFilter(n, 'tukeyIso', [], gpu, gibbs) applied in place to compact
    rows y [Ns, n3, C] with phase-encode indices k1, k2 (FFT order) of the
    n1 x n2 x n3 grid (FFT, window, inverse FFT along the readout)."""
    ctx = ctx or default_context(y.device.index or 0)
    _ptr(y)
    k1 = np.ascontiguousarray(k1, dtype=np.int32)
    k2 = np.ascontiguousarray(k2, dtype=np.int32)
    if y.dim() != 3 or len(k1) != y.shape[0] or len(k2) != y.shape[0] or y.shape[1] != n[2]:
        raise ValueError("filter_rows: y must be [Ns, n3, C] with one (k1, k2) per row")
    ctx._prep()
    check(ctx.lib.dis_filter_rows(ctx.h, _ptr(y), int(y.shape[0]), int(n[2]), int(y.shape[2]),
                                  k1.ctypes.data_as(_lib.c_int32_p), k2.ctypes.data_as(_lib.c_int32_p), int(n[0]),
                                  int(n[1]), float(gibbs)), ctx.h)
    return y


def coil_precond(S, eps=1e-9, ctx: Context = None):
    """Reconstruction/solveXT.m:231-232: Se = 1/(sum_c |S_c|^2 + eps), float32 [n1,n2,n3]."""
    ctx = ctx or default_context(S.device.index or 0)
    nv, C = _coil_view(S, "S")
    Se = torch.empty(tuple(S.shape[:-1])[::-1], dtype=torch.float32, device=S.device).permute(
        *range(S.dim() - 2, -1, -1))
    ctx._prep()
    check(ctx.lib.dis_coil_precond(ctx.h, _ptr(S), nv, C, float(eps), _fptr(Se)), ctx.h)
    return Se
