"""Pins of the oracle's one-time RI preparation (metric contraction, canonical orthogonalization,
AO-leg transform): numpy restatements of the definitions and the invariants that make them unique
(X^T S X = I, X X^T = the kept-space inverse of S, the reconstructed (pq|rs) = B^T B)."""
import numpy as np

from oracle import oracle


def spd(n, seed, drop=0):
    g = np.random.default_rng(seed)
    q, _ = np.linalg.qr(g.standard_normal((n, n)))
    w = np.exp(g.uniform(-3, 2, n))
    w[:drop] = 1e-14  # near-linear dependence
    return (q * w) @ q.T


def test_canonical_orthogonalization_full_rank_is_symmetric_inverse_sqrt():
    s = spd(12, 1)
    x = oracle.canonical_orthogonalize(s, 1e-10)
    assert x.shape == (12, 12)
    assert np.allclose(x, x.T, atol=1e-12)
    assert np.allclose(x @ s @ x, np.eye(12), atol=1e-10)


def test_canonical_orthogonalization_drops_small_eigenvalues():
    s = spd(12, 2, drop=3)
    x = oracle.canonical_orthogonalize(s, 1e-10)
    assert x.shape == (12, 9)
    assert np.allclose(x.T @ s @ x, np.eye(9), atol=1e-10)
    w, v = np.linalg.eigh(s)
    keep = w > 1e-10
    assert np.allclose(x @ x.T, (v[:, keep] / w[keep]) @ v[:, keep].T, rtol=1e-9, atol=1e-8)


def test_metric_contraction_and_ao_legs_match_definitions():
    g = np.random.default_rng(3)
    n, naux_raw = 7, 30
    raw = g.standard_normal((naux_raw, n, n))
    raw = raw + raw.transpose(0, 2, 1)
    metric = spd(naux_raw, 4, drop=2)
    x = oracle.canonical_orthogonalize(metric, 1e-10)
    b = oracle.ri_three_center_tensor(raw.reshape(naux_raw, -1), x)
    assert np.allclose(b, x.T @ raw.reshape(naux_raw, -1), atol=1e-12)
    # the RI integrals do not depend on the (non-unique) choice of X inside the kept space
    w, v = np.linalg.eigh(metric)
    keep = w > 1e-10
    pinv = (v[:, keep] / w[keep]) @ v[:, keep].T
    r = raw.reshape(naux_raw, -1)
    assert np.allclose(b.T @ b, r.T @ pinv @ r, rtol=1e-8, atol=1e-8)
    t = g.standard_normal((n, 5))
    out = oracle.transform_ao_legs(b, n, t)
    want = np.einsum("ao,Pab,bq->Poq", t, b.reshape(-1, n, n), t).reshape(b.shape[0], -1)
    assert np.allclose(out, want, atol=1e-12)
