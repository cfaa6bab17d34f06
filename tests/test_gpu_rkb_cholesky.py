"""GPU parity of the 4c RKB Cholesky path (Fock build, MO transform with and without the
recompression) against the oracle; FP64 complex, 1e-11 relative."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def rkb_case(seed, nl, nchol):
    g = np.random.default_rng(seed)
    cx = lambda *s: g.standard_normal(s) + 1j * g.standard_normal(s)  # noqa: E731
    return g, g.standard_normal((nchol, nl, nl)), cx(nchol, nl, nl), cx(nchol, nl, nl), cx(nchol, nl, nl), cx(nchol, nl, nl)


@pytest.mark.parametrize("nl,nchol", [(5, 11), (13, 60), (30, 150)])
def test_rkb_fock_cholesky(ctx, nl, nchol):
    g, large, s0, s1, cf, cb = rkb_case(nl, nl, nchol)
    n = 4 * nl
    h = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    p = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    p = 0.5 * (p + p.conj().T)
    ctx.rkb_cholesky_upload(0, large, s0, s1, cf, cb)
    f = ctx.rkb_fock_cholesky(0, h, p)
    assert rel(f, oracle.rkb_fock_cholesky(large, s0, s1, cf, cb, h, p)) < 1e-11
    ctx.rkb_cholesky_release(0)


def test_rkb_fock_cholesky_dimension_error(ctx):
    g, large, s0, s1, cf, cb = rkb_case(1, 4, 3)
    ctx.rkb_cholesky_upload(1, large, s0, s1, cf, cb)
    import rerdmft_amd as rr
    with pytest.raises(rr.GpuError, match="multiple of 4"):
        ctx.rkb_fock_cholesky(1, np.zeros((6, 6), complex), np.zeros((6, 6), complex))
    with pytest.raises(rr.GpuError, match="does not match H_RKB"):
        ctx.rkb_fock_cholesky(1, np.zeros((20, 20), complex), np.zeros((20, 20), complex))


@pytest.mark.parametrize("nl,nchol,nmo,nn", [(6, 20, 14, 0), (9, 40, 24, 12), (4, 25, 9, 3)])
def test_rkb_cholesky_to_mo(ctx, nl, nchol, nmo, nn):
    g, large, s0, s1, cf, cb = rkb_case(100 + nl, nl, nchol)
    c = g.standard_normal((4 * nl, nmo)) + 1j * g.standard_normal((4 * nl, nmo))
    ctx.rkb_cholesky_upload(2, large, s0, s1, cf, cb)
    w = ctx.rkb_cholesky_to_mo(2, c, nchol, nn, 1e-12)
    wr = oracle.rkb_cholesky_to_mo(large, s0, s1, cf, cb, c, nn, 1e-12)
    assert w.shape == wr.shape
    if nn == 0:
        assert rel(w, wr) < 1e-11
    else:
        # the recompressed vectors are unique up to the eigenvector phases: compare the invariant
        m = nmo - nn
        a = w.reshape(-1, nmo, nmo)[:, nn:, nn:].reshape(w.shape[0], m * m)
        ar = wr.reshape(-1, nmo, nmo)[:, nn:, nn:].reshape(wr.shape[0], m * m)
        assert rel(a.T @ a.conj(), ar.T @ ar.conj()) < 1e-10
        assert np.all(w.reshape(-1, nmo, nmo)[:, :nn, :] == 0) and np.all(w.reshape(-1, nmo, nmo)[:, :, :nn] == 0)
