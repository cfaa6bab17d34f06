"""GPU parity of the one-time RI preparation (metric contraction, canonical orthogonalization,
AO-leg transform) against the oracle. FP64; 1e-12 relative."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def spd(n, seed, drop=0):
    g = np.random.default_rng(seed)
    q, _ = np.linalg.qr(g.standard_normal((n, n)))
    w = np.exp(g.uniform(-3, 2, n))
    w[:drop] = 1e-14
    return (q * w) @ q.T


@pytest.mark.parametrize("n,naux_raw,drop", [(9, 40, 0), (23, 130, 5), (40, 350, 12)])
def test_metric_pipeline(ctx, n, naux_raw, drop):
    g = np.random.default_rng(n)
    raw = g.standard_normal((naux_raw, n, n))
    raw = (raw + raw.transpose(0, 2, 1)).reshape(naux_raw, -1)
    metric = spd(naux_raw, n + 1, drop)
    x = ctx.canonical_orthogonalize(metric, 1e-10)
    xr = oracle.canonical_orthogonalize(metric, 1e-10)
    assert x.shape == xr.shape == (naux_raw, naux_raw - drop)
    assert np.allclose(x.T @ metric @ x, np.eye(x.shape[1]), atol=1e-9)
    assert rel(x @ x.T, xr @ xr.T) < 1e-9  # unique even where the eigenvector signs are not
    if drop == 0:
        assert rel(x, xr) < 1e-10            # the symmetric inverse square root is unique
    b = ctx.ri_three_center_tensor(raw, xr)
    assert rel(b, oracle.ri_three_center_tensor(raw, xr)) < 1e-12


@pytest.mark.parametrize("naux,n_ao,n_out", [(17, 10, 7), (300, 35, 29), (64, 50, 50)])
def test_ao_legs(ctx, naux, n_ao, n_out):
    g = np.random.default_rng(naux)
    b = g.standard_normal((naux, n_ao * n_ao))
    t = g.standard_normal((n_ao, n_out))
    assert rel(ctx.transform_ri_three_center_ao_legs(b, n_ao, t), oracle.transform_ao_legs(b, n_ao, t)) < 1e-12
