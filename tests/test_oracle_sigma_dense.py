"""Pins the Hartree/exchange sigma oracle (oracle/rr_oracle_sigma.cpp, the RI GEMM form of
src/Hessian_opt/HartreeExchangeSigmaHessianVectorRi.cpp:220-539) where the finite-difference energy
Hessian does not apply: two_rdm_h != two_rdm_x, L1/L2 pair terms, a non-symmetric B and a Fock
matrix that is a free input.

Two dense four-index forms, both built on eri(a,b,c,d) = sum_P B(P,a,c) B(P,b,d):

1. the generalized orbital Hessian's "boxed" Fock-matrix form G_pq,rs
   (src/Hessian_opt/GeneralizedHessian.cpp:36-92) evaluated on a DENSE two-body density matrix that
   holds the Hartree/exchange ansatz Gamma(A,B,C,D) = 1/2 [h(A,B) d_AC d_BD - x(A,B) d_AD d_BC] plus
   the pair terms Gamma(A,pA,C,pC) += l1(A,C) + l2(A,pC)
   (src/Hessian_opt/HartreeExchangeHessian.cpp:22-54, :93-123), with D = diag(occ);
2. the collapsed element G_pq,rs of src/Hessian_opt/HartreeExchangeHessian.cpp:57-91 (H/X) and
   :125-170 (L1/L2), the element the reference's generic dense-ERI sigma
   (src/Hessian_opt/HartreeExchangeSigmaHessianVector.cpp:278-356) contracts.

Phi(a,b) = sum_rs G(a,b,r,s) K(r,s), Psi(c,d) = sum_ab G(a,b,c,d) K(a,b), and
w_pq = 1/2 [(Phi - Phi^T) + (Psi - Psi^T)]_pq (the joint complex form combines the t and y parts as
HartreeExchangeSigmaHessianVector.cpp:339-346 does). The inputs are the reference's own test inputs
(tests/test_hartreeexchange_sigma_hessian_ri.cpp:40-85, seed 864213579, n = 8, n_aux = 11, adjacent
pair involution; and the _joint variant, seed 24681012141618), compared at the reference's
tolerance 1e-9 * max(1, max|w|). Planted mistakes in the oracle (an L2 index transposed, the
exchange coupling used for the Hartree one) are rejected.
"""
import numpy as np
import pytest

from oracle import oracle

from _lcg import Lcg


def reference_inputs(cplx):
    """The draws of tests/test_hartreeexchange_sigma_hessian_ri(_joint).cpp, in their order."""
    rng = Lcg(24681012141618 if cplx else 864213579)
    n, naux = 8, 11
    b = np.zeros((naux, n, n), dtype=np.complex128 if cplx else np.float64)
    for P in range(naux):
        for p in range(n):
            for q in range(n):
                b[P, p, q] = complex(rng.next(), rng.next()) if cplx else rng.next()
    mats = {k: np.zeros((n, n), dtype=np.complex128 if (cplx and k in ("h", "fock")) else np.float64)
            for k in ("h", "fock", "hc", "xc", "l1", "l2")}
    for i in range(n):
        for j in range(n):
            for k in ("h", "fock", "hc", "xc", "l1", "l2"):
                mats[k][i, j] = complex(rng.next(), rng.next()) if (cplx and k in ("h", "fock")) else rng.next()
    occ = np.array([0.1 + 0.8 * (rng.next() + 0.5) for _ in range(n)])
    pair = np.array([i + 1 if i % 2 == 0 else i - 1 for i in range(n)])
    pairs = [(p, q) for p in range(n) for q in range(p)]
    v = np.array([rng.next() for _ in range(len(pairs) * (2 if cplx else 1))])
    return b, mats, occ, pair, pairs, v


def eri(b):
    return np.einsum("Pac,Pbd->abcd", b, b)


def ansatz_gamma(hc, xc, pair=None, l1=None, l2=None):
    n = hc.shape[0]
    d = np.eye(n)
    gam = 0.5 * (np.einsum("ab,ac,bd->abcd", hc, d, d) - np.einsum("ab,ad,bc->abcd", xc, d, d))
    if pair is not None:
        for a in range(n):
            for c in range(n):
                gam[a, pair[a], c, pair[c]] += l1[a, c] + l2[a, pair[c]]
    return gam


def g_boxed(h, g, occ, gam, fock):
    """G[p,q,r,s] of the boxed formula (GeneralizedHessian.cpp:36-92) with D = diag(occ)."""
    n = h.shape[0]
    d1 = np.diag(occ).astype(g.dtype)
    ft = fock.T                                          # F_tex(a, b) = fock(b, a)
    eye = np.eye(n)
    G = np.einsum("qr,sp->pqrs", eye, ft) + np.einsum("sp,rq->pqrs", eye, np.conj(ft))
    G = G - np.einsum("sp,rq->pqrs", h, d1) - np.einsum("qr,ps->pqrs", h, d1)
    G = G - 2 * np.einsum("uvpr,uvsq->pqrs", g, gam)
    G = G + 2 * np.einsum("uspt,ruqt->pqrs", g, gam) - 2 * np.einsum("supt,ruqt->pqrs", g, gam)
    G = G + 2 * np.einsum("uqrt,pust->pqrs", g, gam) - 2 * np.einsum("uqwr,pusw->pqrs", g, gam)
    G = G + 2 * np.einsum("sqwt,rpwt->pqrs", g, gam)
    return G


def g_collapsed(h, g, occ, hc, xc, fock, pair=None, l1=None, l2=None):
    """G[p,q,r,s] of HartreeExchangeHessian.cpp:57-91 (+ :125-170), one element at a time."""
    n = h.shape[0]
    G = np.zeros((n,) * 4, dtype=g.dtype)
    u = np.arange(n)
    for p in range(n):
        for q in range(n):
            for r in range(n):
                for s in range(n):
                    t = 0j if np.iscomplexobj(g) else 0.0
                    if q == r:
                        t += fock[p, s] - h[s, p] * occ[r] + np.sum(hc[r, u] * (g[u, s, p, u] - g[s, u, p, u]))
                    if s == p:
                        t += np.conj(fock[q, r]) - h[q, r] * occ[p] + np.sum(hc[p, u] * (g[u, q, r, u] - g[u, q, u, r]))
                    t += g[s, q, p, r] * (xc[r, q] - hc[s, q] + xc[p, s] - xc[r, p])
                    t += g[s, q, r, p] * (hc[r, p] + xc[q, s] - xc[p, s] - xc[r, q])
                    if pair is not None:
                        if q == pair[s]:
                            t -= 2 * np.sum(g[u, pair[u], p, r] * (l1[u, s] + l2[u, q]))
                        c = l1[r, q] + l2[r, pair[q]]
                        t += 2 * c * (g[pair[r], s, p, pair[q]] - g[s, pair[r], p, pair[q]])
                        c = l1[p, s] + l2[p, pair[s]]
                        t += 2 * c * (g[pair[p], q, r, pair[s]] - g[pair[p], q, pair[s], r])
                        if p == pair[r]:
                            t += 2 * np.sum(g[s, q, u, pair[u]] * (l1[r, u] + l2[r, pair[u]]))
                    G[p, q, r, s] = t
    return G


def sigma_from_g(G, pairs, v, joint):
    n = G.shape[0]
    npairs = len(pairs)
    kt = np.zeros((n, n))
    ky = np.zeros((n, n))
    for i, (p, q) in enumerate(pairs):
        kt[p, q], kt[q, p] = v[i], -v[i]
        if joint:
            ky[p, q] = ky[q, p] = v[npairs + i]
    phi_t, psi_t = np.einsum("abrs,rs->ab", G, kt), np.einsum("abcd,ab->cd", G, kt)
    if not joint:
        return np.array([0.5 * ((phi_t[p, q] - phi_t[q, p]) + (psi_t[p, q] - psi_t[q, p])) for p, q in pairs])
    phi_y, psi_y = np.einsum("abrs,rs->ab", G, ky), np.einsum("abcd,ab->cd", G, ky)
    w = np.zeros(2 * npairs)
    for i, (p, q) in enumerate(pairs):
        w[i] = np.real(0.5 * ((phi_t[p, q] - phi_t[q, p]) + (psi_t[p, q] - psi_t[q, p]))
                       + 0.5j * ((phi_y[p, q] - phi_y[q, p]) + (psi_y[p, q] - psi_y[q, p])))
        w[npairs + i] = np.real(0.5j * ((psi_t[p, q] + psi_t[q, p]) + (phi_t[p, q] + phi_t[q, p]))
                                - 0.5 * ((phi_y[p, q] + phi_y[q, p]) + (psi_y[p, q] + psi_y[q, p])))
    return w


def oracle_sigma(b, m, occ, pair, pairs, v, with_l=True):
    naux = b.shape[0]
    kw = dict(pair_of=pair, l1=m["l1"], l2=m["l2"]) if with_l else {}
    return oracle.sigma_hessian_vector_ri(m["h"], b.reshape(naux, -1), occ, m["hc"], m["xc"], pairs, v,
                                          fock=m["fock"], hx=True, **kw)


@pytest.mark.parametrize("cplx", [False, True])
def test_boxed_and_collapsed_hessian_elements_agree(cplx):
    """The collapse of the six two-body sums (HartreeExchangeHessian.cpp:22-54, :93-123) holds."""
    b, m, occ, pair, _, _ = reference_inputs(cplx)
    g = eri(b)
    gam = ansatz_gamma(m["hc"], m["xc"], pair, m["l1"], m["l2"])
    gb = g_boxed(m["h"], g, occ, gam, m["fock"])
    gc = g_collapsed(m["h"], g, occ, m["hc"], m["xc"], m["fock"], pair, m["l1"], m["l2"])
    assert np.max(np.abs(gb - gc)) < 1e-12 * max(1.0, np.max(np.abs(gb)))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("with_l", [True, False])
def test_ri_sigma_oracle_matches_dense_forms(cplx, with_l):
    b, m, occ, pair, pairs, v = reference_inputs(cplx)
    g = eri(b)
    if with_l:
        gam = ansatz_gamma(m["hc"], m["xc"], pair, m["l1"], m["l2"])
    else:
        gam = ansatz_gamma(m["hc"], m["xc"])
    w_boxed = sigma_from_g(g_boxed(m["h"], g, occ, gam, m["fock"]), pairs, v, cplx)
    kw = dict(pair=pair, l1=m["l1"], l2=m["l2"]) if with_l else {}
    w_coll = sigma_from_g(g_collapsed(m["h"], g, occ, m["hc"], m["xc"], m["fock"], **kw), pairs, v, cplx)
    w = oracle_sigma(b, m, occ, pair, pairs, v, with_l)
    tol = 1e-9 * max(1.0, np.max(np.abs(w_boxed)))
    assert np.max(np.abs(w_coll - w_boxed)) < tol
    assert np.max(np.abs(w - w_boxed)) < tol
    # the couplings really differ, so the h and x terms are told apart
    assert np.max(np.abs(m["hc"] - m["xc"])) > 0.1


@pytest.mark.parametrize("plant", ["hx_l2_flip", "hx_h_as_x"])
@pytest.mark.parametrize("cplx", [False, True])
def test_planted_sigma_mistakes_are_rejected(monkeypatch, plant, cplx):
    b, m, occ, pair, pairs, v = reference_inputs(cplx)
    gam = ansatz_gamma(m["hc"], m["xc"], pair, m["l1"], m["l2"])
    w_boxed = sigma_from_g(g_boxed(m["h"], eri(b), occ, gam, m["fock"]), pairs, v, cplx)
    monkeypatch.setenv("RR_ORACLE_PLANT", plant)
    w = oracle_sigma(b, m, occ, pair, pairs, v)
    assert np.max(np.abs(w - w_boxed)) > 1e-3 * max(1.0, np.max(np.abs(w_boxed)))
