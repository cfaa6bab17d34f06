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


def hermitian_case(seed, nl, nchol):
    """RKB Cholesky vectors as the reference builds them: large real symmetric, the two small
    (projected) components Hermitian; the cross components enter the Coulomb part only."""
    g, large, s0, s1, cf, cb = rkb_case(seed, nl, nchol)
    large = 0.5 * (large + large.transpose(0, 2, 1))
    s0 = 0.5 * (s0 + s0.conj().transpose(0, 2, 1))
    s1 = 0.5 * (s1 + s1.conj().transpose(0, 2, 1))
    return g, large, s0, s1, cf, cb


def rkb_density(g, n, n_occ, nx, occupations):
    """An S-orthonormal occupied set in the RKB basis (S Hermitian positive definite, X its canonical
    orthogonalizer with nx columns) and P = C diag(occ) C^H."""
    a = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    s = a @ a.conj().T / n + np.eye(n)
    w, v = np.linalg.eigh(s)
    x = (v / np.sqrt(w))[:, n - nx:]
    q, _ = np.linalg.qr(g.standard_normal((nx, n_occ)) + 1j * g.standard_normal((nx, n_occ)))
    c = x @ q
    return s, x, (c * occupations) @ c.conj().T


@pytest.mark.parametrize("nl,nchol,n_occ,frac", [(5, 11, 3, False), (13, 60, 10, False), (30, 150, 24, True),
                                                 (12, 40, 0, False)])
def test_rkb_fock_cholesky_factor_form(ctx, nl, nchol, n_occ, frac):
    """Factor form (K = sum_L W_L W_L^H over the recovered occupied set) against the oracle's density form
    for P = C diag(occ) C^H: idempotent, and occupations in (1/2, 1]."""
    g, large, s0, s1, cf, cb = hermitian_case(300 + nl, nl, nchol)
    n = 4 * nl
    nx = n - 2
    occ = g.uniform(0.55, 1.0, n_occ) if frac else np.ones(n_occ)
    s, x, p = rkb_density(g, n, n_occ, nx, occ)
    h = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    h = 0.5 * (h + h.conj().T)
    ctx.rkb_cholesky_upload(0, large, s0, s1, cf, cb)
    f = ctx.rkb_fock_cholesky_factor(0, h, s, x, p)
    fr = oracle.rkb_fock_cholesky(large, s0, s1, cf, cb, h, p)
    assert rel(f, fr) < 1e-10
    assert rel(f, ctx.rkb_fock_cholesky(0, h, p)) < 1e-10
    ctx.rkb_cholesky_release(0)


def test_rkb_fock_cholesky_factor_form_needs_hermitian_vectors(ctx):
    g, large, s0, s1, cf, cb = rkb_case(7, 6, 9)  # small components not Hermitian
    n = 24
    s, x, p = rkb_density(g, n, 4, n, np.ones(4))
    ctx.rkb_cholesky_upload(3, large, s0, s1, cf, cb)
    import rerdmft_amd as rr
    with pytest.raises(rr.GpuError, match="not Hermitian"):
        ctx.rkb_fock_cholesky_factor(3, np.zeros((n, n), complex), s, x, p)
    ctx.rkb_cholesky_release(3)
