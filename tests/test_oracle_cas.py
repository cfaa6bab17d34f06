"""CPU pins of the CAS oracle (oracle/cas.py) against the dense four-index tensor of the same RI MO
tensor: the two-electron operator (real and complex, definite and indefinite rho), the active
integrals and the active cumulant Fock matrix, with planted mistakes rejected."""
import numpy as np
import pytest

from oracle import cas


def _tensor(naux, n, seed, cplx):
    g = np.random.default_rng(seed)
    b = g.standard_normal((naux, n, n)) + (1j * g.standard_normal((naux, n, n)) if cplx else 0)
    return 0.5 * (b + b.conj().transpose(0, 2, 1))  # Hermitian B_P


def _dense(b):
    # <ab|cd> = sum_P B(P, a n + c) B(P, b n + d)
    return np.einsum("Pac,Pbd->abcd", b, b)


def _rho(n, seed, cplx, rank=None):
    g = np.random.default_rng(seed)
    a = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0)
    r = a + a.conj().T
    if rank is not None:
        w, u = np.linalg.eigh(r)
        r = (u[:, :rank] * w[:rank]) @ u[:, :rank].conj().T
    return r


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("rank", [None, 3])
def test_two_electron_operator_matches_dense_definition(cplx, rank):
    n, naux = 7, 11
    b = _tensor(naux, n, 1, cplx)
    rho = _rho(n, 2, cplx, rank)
    eri = _dense(b)
    # CasFock.h:44-58: G_rp = sum_ab rho_ab (<ar|bp> - <ra|bp>)
    ref = np.einsum("ab,arbp->rp", rho, eri) - np.einsum("ab,rabp->rp", rho, eri)
    got = cas.two_electron_operator(b, n, rho)
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))
    bad = cas.two_electron_operator(b, n, rho, planted=True)
    assert np.max(np.abs(bad - ref)) > 1e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("cplx", [False, True])
def test_active_integrals_match_dense(cplx):
    n, naux = 8, 9
    b = _tensor(naux, n, 3, cplx)
    act = [1, 4, 6]
    eri = _dense(b)
    # (tu|vw) = <tv|uw>
    ref = np.array([[eri[t, v, u, w] for v in act for w in act] for t in act for u in act])
    np.testing.assert_allclose(cas.active_integrals(b, n, act), ref, atol=1e-12)


@pytest.mark.parametrize("cplx", [False, True])
def test_active_cumulant_fock_matches_dense(cplx):
    n, naux = 8, 10
    b = _tensor(naux, n, 5, cplx)
    act = [0, 3, 5]
    a = len(act)
    g = np.random.default_rng(6)
    lam = g.standard_normal((a, a, a, a)) + (1j * g.standard_normal((a, a, a, a)) if cplx else 0)
    eri = _dense(b)
    # F(p,q) = sum_{r,s,t} lambda[s,t,r,q] sum_P B(P, act_r n + p) B(P, act_s n + act_t)
    #        = sum_{r,s,t} lambda[s,t,r,q] <act_r act_s | p act_t>
    ref = np.zeros((n, a), dtype=lam.dtype)
    for p in range(n):
        for q in range(a):
            ref[p, q] = sum(lam[s, t, r, q] * eri[act[r], act[s], p, act[t]]
                            for r in range(a) for s in range(a) for t in range(a))
    got = cas.active_cumulant_fock(b, n, act, lam.reshape(-1))
    np.testing.assert_allclose(got, ref, atol=1e-11)
    bad = cas.active_cumulant_fock(b, n, act, lam.reshape(-1), planted=True)
    assert np.max(np.abs(bad - ref)) > 1e-3


@pytest.mark.parametrize("cplx", [False, True])
def test_physics_block_matches_dense(cplx):
    n, naux = 9, 6
    b = _tensor(naux, n, 11, cplx)
    eri = _dense(b)  # <ab|cd> = sum_P B(P, a n + c) B(P, b n + d)
    p, q, r, s = [0, 4], [1, 2, 8], [3], [5, 6]
    ref = eri[np.ix_(p, q, r, s)]
    np.testing.assert_allclose(cas.physics_block(b, n, p, q, r, s), ref, atol=1e-12)
    # slab(p, list) is the block ({p}, list, list, list)
    lst = [1, 3, 7]
    np.testing.assert_allclose(cas.physics_block(b, n, [2], lst, lst, lst)[0], eri[2][np.ix_(lst, lst, lst)], atol=1e-12)
