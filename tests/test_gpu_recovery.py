"""Occupied-set recovery of the density-taking builds. The reference diagonalizes X^H (S P S) X and
keeps the eigenvectors above 1/2 (NonRelFockMatrixRi.cpp:51-69, X2C_FockMatrixRi.cpp:65-95,
UkbFockMatrixRi.cpp:185-211); the library's default instead builds the spectral projector onto those
eigenvalues (Newton-Schulz sign iteration) and a factor of the occupied part. Both methods must give
the oracle's Fock matrix for idempotent and fractional densities; an eigenvalue at 1/2 must send the
build to the eigensolver; no eigenvalue above 1/2 leaves no exchange."""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn

pytestmark = pytest.mark.gpu

SLOT = 9


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture
def method(ctx):
    yield
    ctx.recovery_method(0)


def _run_both(ctx, build):
    """Fock matrix by the projector and by the eigensolver, with the counters' increments."""
    ctx.recovery_method(0)
    a0, e0 = ctx.recovery_method()
    f0 = build()
    a1, e1 = ctx.recovery_method()
    ctx.recovery_method(1)
    f1 = build()
    a2, e2 = ctx.recovery_method()
    assert (a2, e2 - e1) == (a1, 1)
    return f0, f1, (a1 - a0, e1 - e0)


def _nonrel(n, naux, weights, seed):
    b = syn.ri_tensor(naux, n, 0.3, 6.0, seed)
    s = syn.spd_overlap(n, seed + 1)
    x = syn.inverse_sqrt(s)
    c = syn.s_orthonormal_occ(s, len(weights), seed + 2)
    p = 2.0 * (c * np.asarray(weights)) @ c.T  # eigenvalues of X^T (S P/2 S) X = weights
    return b, s, x, p, syn.symmetric(n, seed + 3)


@pytest.mark.parametrize("weights", [[1.0] * 9, [0.95, 0.6, 0.2, 1.0, 0.51, 0.49, 0.999, 0.05]])
def test_nonrel_projector_matches_eigensolver_and_oracle(ctx, oracle, method, weights):
    n, naux = 48, 120
    b, s, x, p, h = _nonrel(n, naux, weights, 611)
    ctx.upload(SLOT, b)
    fr = oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p)
    f0, f1, used = _run_both(ctx, lambda: ctx.nonrel_fock_ri(SLOT, h, x, s, p))
    assert used == (1, 0)
    assert rel(f0, fr) < 1e-12 and rel(f1, fr) < 1e-12


def test_nonrel_full_and_empty_occupation(ctx, oracle, method):
    n, naux = 16, 40
    b, s, x, p, h = _nonrel(n, naux, [1.0] * n, 621)  # every eigenvalue 1: r = n
    ctx.upload(SLOT, b)
    f0, f1, used = _run_both(ctx, lambda: ctx.nonrel_fock_ri(SLOT, h, x, s, p))
    assert used == (1, 0)
    fr = oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p)
    assert rel(f0, fr) < 1e-12 and rel(f1, fr) < 1e-12
    p_low = 0.6 * p  # every eigenvalue 0.6: still occupied
    f0 = ctx.nonrel_fock_ri(SLOT, h, x, s, p_low)
    assert rel(f0, oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p_low)) < 1e-12
    p_none = 0.4 * p  # every eigenvalue 0.4: no exchange, F = h + J
    f0, j0, k0 = ctx.nonrel_fock_ri(SLOT, h, x, s, p_none, return_jk=True)
    assert np.all(k0 == 0.0)
    assert rel(f0, oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p_none)) < 1e-12


def test_eigenvalue_at_one_half_uses_eigensolver(ctx, oracle, method):
    n, naux = 30, 70
    b, s, x, p, h = _nonrel(n, naux, [1.0, 0.5 + 1e-13, 0.3, 0.8], 631)
    ctx.upload(SLOT, b)
    f0, f1, used = _run_both(ctx, lambda: ctx.nonrel_fock_ri(SLOT, h, x, s, p))
    assert used == (0, 1)  # the sign iteration cannot separate 1/2 + 1e-13 in its step budget
    fr = oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p)
    assert rel(f0, fr) < 1e-12 and np.array_equal(f0, f1)


@pytest.mark.parametrize("weights", [None, [1.0, 0.97, 0.7, 0.55, 0.45, 0.9, 0.2, 1.0]])
def test_x2c_projector_matches_eigensolver_and_oracle(ctx, oracle, method, weights):
    nl, naux = 40, 100
    b = syn.ri_tensor(naux, nl, 0.5, nl / 4.0, 641)
    s1 = syn.spd_overlap(nl, 642)
    s = np.kron(np.eye(2), s1).astype(np.complex128)
    x = np.kron(np.eye(2), syn.inverse_sqrt(s1)).astype(np.complex128)
    if weights is None:
        q = syn.kramers_pair_occ(nl, 6, 643)
        w = np.ones(q.shape[1])
    else:
        q = syn.unitary(2 * nl, 643)[:, :len(weights)]
        w = np.asarray(weights)
    c = x @ q
    p = (c * w) @ c.conj().T
    h = syn.hermitian(2 * nl, 644)
    ctx.upload(SLOT, b)
    fr = oracle.x2c_fock_ri(b.reshape(naux, -1), x, s, h, p)
    f0, f1, used = _run_both(ctx, lambda: ctx.x2c_fock_ri(SLOT, h, x, s, p))
    assert used == (1, 0)
    assert rel(f0, fr) < 1e-11 and rel(f1, fr) < 1e-11
    assert np.max(np.abs(f0 - f0.conj().T)) < 1e-12 * np.max(np.abs(f0))


def test_rkb_projector_matches_eigensolver_and_oracle(ctx, oracle, method):
    nl, ns, naux, nocc = 9, 14, 30, 6
    bll = syn.ri_tensor(naux, nl, 0.4, 6.0, 651)
    bss = syn.ri_tensor(naux, ns, 0.2, 6.0, 652)
    n_ukb, n_rkb = 2 * nl + 2 * ns, 4 * nl
    g = syn.rng(653)
    v = (g.standard_normal((n_ukb, n_rkb)) + 1j * g.standard_normal((n_ukb, n_rkb))) * 0.3
    sr = v.conj().T @ v + 0.5 * np.eye(n_rkb)
    w, u = np.linalg.eigh(sr)
    x = u / np.sqrt(w)[None, :]
    c = x @ syn.unitary(n_rkb, 654)[:, :nocc]
    p = (c * np.array([1.0, 0.9, 0.75, 1.0, 0.3, 0.6])) @ c.conj().T
    h = syn.hermitian(n_rkb, 655)
    ctx.upload(SLOT, bll)
    ctx.upload(SLOT - 1, bss)
    fr = oracle.rkb_fock_ri(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), v, sr, x, h, p)
    f0, f1, used = _run_both(ctx, lambda: ctx.rkb_fock_ri(SLOT, SLOT - 1, nl, ns, h, v, sr, x, p))
    assert used == (1, 0)
    assert rel(f0, fr) < 1e-11 and rel(f1, fr) < 1e-11


def test_recovery_method_rejects_unknown_value(ctx):
    with pytest.raises(rr.GpuError):
        ctx.recovery_method(2)


@pytest.mark.parametrize("nocc", [170, 230])
def test_nonrel_large_occupied_set(ctx, oracle, method, nocc):
    n, naux = 240, 40
    weights = [1.0] * (nocc - 10) + [0.9, 0.8, 0.7, 0.6, 0.55, 0.95, 0.85, 0.75, 0.65, 0.52]
    b, s, x, p, h = _nonrel(n, naux, weights, 671)
    ctx.upload(SLOT, b)
    fr = oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p)
    f0, f1, used = _run_both(ctx, lambda: ctx.nonrel_fock_ri(SLOT, h, x, s, p))
    assert used == (1, 0)
    assert rel(f0, fr) < 1e-11 and rel(f1, fr) < 1e-11


@pytest.mark.parametrize("nocc", [190, 215])
def test_x2c_large_occupied_set(ctx, oracle, method, nocc):
    nl, naux = 110, 30
    b = syn.ri_tensor(naux, nl, 0.5, nl / 4.0, 681)
    s1 = syn.spd_overlap(nl, 682)
    s = np.kron(np.eye(2), s1).astype(np.complex128)
    x = np.kron(np.eye(2), syn.inverse_sqrt(s1)).astype(np.complex128)
    q = syn.unitary(2 * nl, 683)[:, :nocc]
    w = np.linspace(1.0, 0.55, nocc)
    c = x @ q
    p = (c * w) @ c.conj().T
    h = syn.hermitian(2 * nl, 684)
    ctx.upload(SLOT, b)
    fr = oracle.x2c_fock_ri(b.reshape(naux, -1), x, s, h, p)
    f0, f1, used = _run_both(ctx, lambda: ctx.x2c_fock_ri(SLOT, h, x, s, p))
    assert used == (1, 0)
    assert rel(f0, fr) < 1e-11 and rel(f1, fr) < 1e-11
