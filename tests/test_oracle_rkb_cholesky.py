"""Pins of the oracle's 4c RKB Cholesky restatement against dense 4-component contractions.

With the full RKB-space vector Bfull_L = blockdiag(large, large, small0, small1) plus the cross
blocks small_cross[0] at (Sa, Sb) and small_cross[1] at (Sb, Sa):
  J = sum_L tr(Bfull_L P) Bfull_L          (the Coulomb part including the cross-flavor terms)
  K = sum_L Bdiag_L P Bdiag_L              (exchange over the flavor-diagonal blocks only)
  W_L = (C^H Bfull_L C)[pos, pos]^T        (the MO vectors; pos = the positive-energy columns)
and the recompressed vectors reproduce sum_L W_L (x) conj(W_L) (the Gram-matrix invariant)."""
import numpy as np

from oracle import oracle


def rkb_case(seed, nl, nchol):
    g = np.random.default_rng(seed)
    cx = lambda *s: g.standard_normal(s) + 1j * g.standard_normal(s)  # noqa: E731
    large = g.standard_normal((nchol, nl, nl))
    return g, large, cx(nchol, nl, nl), cx(nchol, nl, nl), cx(nchol, nl, nl), cx(nchol, nl, nl)


def full_vectors(nl, large, s0, s1, cf, cb, cross=True):
    nchol = large.shape[0]
    b = np.zeros((nchol, 4 * nl, 4 * nl), dtype=complex)
    for f, blk in enumerate([large, large, s0, s1]):
        b[:, f * nl:(f + 1) * nl, f * nl:(f + 1) * nl] = blk
    if cross:
        b[:, 2 * nl:3 * nl, 3 * nl:] = cf
        b[:, 3 * nl:, 2 * nl:3 * nl] = cb
    return b


def test_rkb_cholesky_fock_matches_dense():
    nl, nchol = 4, 7
    g, large, s0, s1, cf, cb = rkb_case(5, nl, nchol)
    n = 4 * nl
    h = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    p = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    f = oracle.rkb_fock_cholesky(large, s0, s1, cf, cb, h, p)
    bf = full_vectors(nl, large, s0, s1, cf, cb)
    bd = full_vectors(nl, large, s0, s1, cf, cb, cross=False)
    t = np.einsum("Lqs,sq->L", bf, p)
    j = np.einsum("L,Lab->ab", t, bf)
    k = np.einsum("Lam,mn,Lnb->ab", bd, p, bd)
    assert np.allclose(f, h + j - k, atol=1e-11)


def test_rkb_cholesky_to_mo_and_recompression():
    nl, nchol, nmo, nn = 3, 9, 10, 4
    g, large, s0, s1, cf, cb = rkb_case(6, nl, nchol)
    c = g.standard_normal((4 * nl, nmo)) + 1j * g.standard_normal((4 * nl, nmo))
    bf = full_vectors(nl, large, s0, s1, cf, cb)
    w_full = np.einsum("ap,Lab,bq->Lpq", c.conj(), bf, c)
    w = oracle.rkb_cholesky_to_mo(large, s0, s1, cf, cb, c, 0)
    assert np.allclose(w.reshape(nchol, nmo, nmo), np.transpose(w_full, (0, 2, 1)), atol=1e-11)
    wc = oracle.rkb_cholesky_to_mo(large, s0, s1, cf, cb, c, nn, threshold=1e-12)
    a = np.transpose(w_full[:, nn:, nn:], (0, 2, 1)).reshape(nchol, -1)
    ap = wc.reshape(-1, nmo, nmo)[:, nn:, nn:].reshape(wc.shape[0], -1)
    assert np.allclose(wc.reshape(-1, nmo, nmo)[:, :nn, :], 0) and np.allclose(wc.reshape(-1, nmo, nmo)[:, :, :nn], 0)
    assert np.allclose(ap.T @ ap.conj(), a.T @ a.conj(), atol=1e-9)
