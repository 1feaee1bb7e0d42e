"""FP64 CPU restatement of the learned motion-proposal network (test
infrastructure only: imported by tests/, never by the product).

The network has no counterpart in the reference; its definition is the
one stated for the learned proposal: per shot, an MLP 13 -> 256 -> 256 -> 6
with GELU (erf form) and a 1D convolution along shot order (kernel 5,
6 -> 6 channels, zero padding inside each scan).  Parameters are one flat
vector: W1 [256x13], b1, W2 [256x256], b2, W3 [6x256], b3, Wc [6x6x5]
(out, in, tap), bc.

``bf16=True`` rounds both operands of every matrix product to bfloat16
(round to nearest even), as the matrix cores do; the products and sums are
then exact in FP64, which leaves only the device's FP32 accumulation as the
difference.  ``bf16=False`` is the plain FP64 network the finite-difference
checks pin.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

NIN, NHID, NOUT, NK = 13, 256, 6, 5
SHAPES = [("W1", (NHID, NIN)), ("b1", (NHID,)), ("W2", (NHID, NHID)), ("b2", (NHID,)), ("W3", (NOUT, NHID)),
          ("b3", (NOUT,)), ("Wc", (NOUT, NOUT, NK)), ("bc", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16_round(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, bf16):
    if bf16:
        a, b = bf16_round(a), bf16_round(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _conv(U, Wc, bc, scan_off):
    Y = np.tile(bc, (U.shape[0], 1)).astype(np.float64)
    for b in range(len(scan_off) - 1):
        lo, hi = scan_off[b], scan_off[b + 1]
        for r in range(lo, hi):
            for t in range(NK):
                rr = r + t - NK // 2
                if lo <= rr < hi:
                    Y[r] += Wc[:, :, t] @ U[rr]
    return Y


def forward(p, X, scan_off, bf16=True):
    q = unpack(p)
    X = np.asarray(X, np.float64)
    Z1 = _mm(X, q["W1"].T, bf16) + q["b1"]
    H1 = gelu(Z1)
    Z2 = _mm(H1, q["W2"].T, bf16) + q["b2"]
    H2 = gelu(Z2)
    U = _mm(H2, q["W3"].T, bf16) + q["b3"]
    Y = _conv(U, q["Wc"], q["bc"], scan_off)
    return Y, (X, Z1, H1, Z2, H2, U)


def backward(p, X, scan_off, dY, bf16=True):
    """Gradients of sum(dY * Y) w.r.t. the flat parameters and X."""
    q = unpack(p)
    Y, (X, Z1, H1, Z2, H2, U) = forward(p, X, scan_off, bf16)
    dY = np.asarray(dY, np.float64)
    g = {}
    dU = np.zeros_like(U)
    g["Wc"] = np.zeros_like(q["Wc"])
    g["bc"] = dY.sum(axis=0)
    for b in range(len(scan_off) - 1):
        lo, hi = scan_off[b], scan_off[b + 1]
        for r in range(lo, hi):
            for t in range(NK):
                rr = r + t - NK // 2
                if lo <= rr < hi:
                    g["Wc"][:, :, t] += np.outer(dY[r], U[rr])
                    dU[rr] += q["Wc"][:, :, t].T @ dY[r]
    g["W3"] = _mm(dU.T, H2, bf16)
    g["b3"] = dU.sum(axis=0)
    dH2 = _mm(dU, q["W3"], bf16)
    dZ2 = dH2 * gelu_grad(Z2)
    g["W2"] = _mm(dZ2.T, H1, bf16)
    g["b2"] = dZ2.sum(axis=0)
    dH1 = _mm(dZ2, q["W2"], bf16)
    dZ1 = dH1 * gelu_grad(Z1)
    g["W1"] = _mm(dZ1.T, X, bf16)
    g["b1"] = dZ1.sum(axis=0)
    dX = _mm(dZ1, q["W1"], bf16)
    return pack(g), dX


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
