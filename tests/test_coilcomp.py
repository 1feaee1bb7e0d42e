"""Coil-array compression (Methods/compressCoils.m:75-118).

CPU: the host eigensolver behind dis_coil_basis against numpy.linalg.eigh and
the oracle's SVD-based compress_coils (P is Hermitian PSD, so svd(P) is its
eigendecomposition; column phases are arbitrary, so subspaces are compared
through their projectors U U^H).
GPU: the Gram kernel against the FP64 oracle (rel-L2 <= 1e-6), the MFMA
split-bf16 contraction against an FP64 product (terms=3: rel-L2 <= 3e-5;
terms=6: <= 1e-6), and the whole compressCoils through phase-invariant
quantities (D, component counts, projected maps)."""
import numpy as np
import pytest

from oracle import disorder_oracle as O
from tests.helpers import crandn, rel_l2


def _basis(P, perc):
    import disorder_amd as D
    return D.coil_basis(P, perc)


def _random_psd(rng, C, rank=None, decay=0.7):
    rank = rank or C
    B = crandn(rng, (C, rank)) * (decay ** np.arange(rank))[None, :]
    return B @ np.conj(B).T


@pytest.mark.parametrize("C", [1, 2, 8, 32, 64])
def test_basis_matches_eigh(C):
    rng = np.random.default_rng(100 + C)
    P = _random_psd(rng, C)
    U, Dv, cc = _basis(P, [0.9, 0.99, 3])
    w = np.linalg.eigvalsh(P)[::-1]
    assert np.allclose(Dv, np.abs(w), rtol=1e-10, atol=1e-12 * w[0])
    assert np.allclose(np.conj(U).T @ U, np.eye(C), atol=1e-12)
    assert np.allclose(U @ np.diag(Dv) @ np.conj(U).T, P, atol=1e-11 * np.abs(P).max())
    cs = np.cumsum(Dv) / Dv.sum()
    for p, k in zip([0.9, 0.99], cc[:2]):
        assert k == int(np.argmax(cs >= p) + 1)
    assert cc[2] == min(3, C)


def test_basis_projectors_match_oracle_svd():
    rng = np.random.default_rng(7)
    n, C = (6, 5, 4), 12
    Sm = crandn(rng, n + (4,)) @ crandn(rng, (4, C)) + 1e-2 * crandn(rng, n + (C,))
    _, _, P, A = O.compress_coils(Sm, 0.99)
    U, Dv, cc = _basis(P, 0.99)
    k = int(cc[0])
    assert k == A.shape[0]
    Pg = U[:, :k] @ np.conj(U[:, :k]).T
    Po = np.conj(A).T @ A
    assert np.allclose(Pg, Po, atol=1e-10)
    assert np.allclose(Dv, np.linalg.svd(P, compute_uv=False), rtol=1e-10, atol=1e-14)


def test_basis_repeated_eigenvalues_and_zero():
    # identity (every eigenvalue equal) and the zero matrix
    U, Dv, cc = _basis(np.eye(4, dtype=np.complex128) * 2.0, 0.5)
    assert np.allclose(Dv, 2.0) and cc[0] == 2
    U, Dv, cc = _basis(np.zeros((3, 3), np.complex128), 2)
    assert np.allclose(Dv, 0.0) and cc[0] == 2
    assert np.allclose(U, np.eye(3))


# ---------------------------------------------------------------- GPU
def _to_dev(a, dev):
    import disorder_amd as D
    import torch
    t = torch.from_numpy(np.asfortranarray(np.asarray(a, np.complex64)))
    return D.colmajor(t.to(dev))


def _maps(rng, n, C, rank=6):
    """Smooth low-rank-ish coil maps: a few blobs mixed into C channels."""
    g = np.meshgrid(*[np.linspace(-1, 1, v) for v in n], indexing="ij")
    base = []
    for _ in range(rank):
        c = rng.uniform(-0.6, 0.6, 3)
        r2 = sum((gi - ci) ** 2 for gi, ci in zip(g, c))
        base.append(np.exp(-r2 / 0.5) * np.exp(1j * (rng.uniform(-2, 2) * g[0] + rng.uniform(-2, 2) * g[1])))
    base = np.stack(base, -1)
    return base @ crandn(rng, (rank, C)) + 1e-3 * crandn(rng, n + (C,))


@pytest.mark.gpu
@pytest.mark.parametrize("n,C", [((16, 12, 10), 8), ((20, 18, 16), 32), ((24, 16, 12), 64), ((7, 5, 3), 3)])
def test_gram_matches_oracle(dev, n, C):
    import disorder_amd as D
    rng = np.random.default_rng(C + sum(n))
    Sm = _maps(rng, n, C)
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    _, _, Pref, _ = O.compress_coils(Sm32, C)
    P = D.coil_gram(_to_dev(Sm, dev))
    assert rel_l2(P, Pref) < 1e-6
    assert np.allclose(P, np.conj(P).T, atol=1e-6 * np.abs(P).max())


@pytest.mark.gpu
@pytest.mark.parametrize("terms,tol", [(3, 3e-5), (6, 1e-6)])
@pytest.mark.parametrize("nv,C,Cc", [(4096, 64, 8), (1000, 32, 8), (333, 8, 3), (130, 64, 20), (64, 5, 5)])
def test_apply_matches_fp64(dev, terms, tol, nv, C, Cc):
    import disorder_amd as D
    rng = np.random.default_rng(nv + C + Cc)
    X = crandn(rng, (nv, C)).astype(np.complex64)
    A = crandn(rng, (Cc, C))
    out = D.coil_apply(A, _to_dev(X, dev), terms=terms).cpu().numpy()
    ref = X.astype(np.complex128) @ A.T
    err = rel_l2(out, ref)
    assert err < tol, (terms, err)


@pytest.mark.gpu
def test_apply_split_terms_order_errors(dev):
    """The six-product split is more accurate than the three-product one."""
    import disorder_amd as D
    rng = np.random.default_rng(5)
    X = crandn(rng, (8192, 64)).astype(np.complex64)
    A = crandn(rng, (8, 64))
    ref = X.astype(np.complex128) @ A.T
    e3 = rel_l2(D.coil_apply(A, _to_dev(X, dev), terms=3).cpu().numpy(), ref)
    e6 = rel_l2(D.coil_apply(A, _to_dev(X, dev), terms=6).cpu().numpy(), ref)
    assert e6 < e3 < 3e-5
    assert e6 < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("terms", [3, 6])
def test_compress_coils_end_to_end(dev, terms):
    import disorder_amd as D
    rng = np.random.default_rng(11)
    n, C = (24, 20, 16), 32
    Sm = _maps(rng, n, C, rank=8)
    ydat = crandn(rng, (300, 16, C))
    Sm32 = Sm.astype(np.complex64).astype(np.complex128)
    y32 = ydat.astype(np.complex64).astype(np.complex128)
    Sc_ref, yc_ref, P, A = O.compress_coils(Sm32, 6, y32, ncomp=6)
    Sc, yc, cc, Dv, U = D.compressCoils(_to_dev(Sm, dev), 6, _to_dev(ydat, dev), terms=terms)
    k = A.shape[0]
    assert int(cc[0]) == k and Sc.shape[-1] == k and yc.shape[-1] == k
    assert np.allclose(Dv, np.linalg.svd(P, compute_uv=False), rtol=1e-5, atol=1e-6 * Dv[0])
    # phase-invariant: maps and data projected back into coil space
    Ag = np.conj(U[:, :k]).T
    back = Sc.cpu().numpy().astype(np.complex128) @ np.conj(Ag)
    back_ref = Sc_ref @ np.conj(A)
    assert rel_l2(back, back_ref) < 3e-5
    yback = yc.cpu().numpy().astype(np.complex128) @ np.conj(Ag)
    assert rel_l2(yback, yc_ref @ np.conj(A)) < 3e-5
    # per-component magnitudes agree (phases are arbitrary)
    assert rel_l2(np.abs(Sc.cpu().numpy()), np.abs(Sc_ref)) < 3e-5
