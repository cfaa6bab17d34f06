"""Pins of the functional-stage oracle (oracle/rr_oracle_sigma.cpp) at tiny sizes.

- riHxFockGemm: the dense definition of the generalized Fock matrix
  (src/Hessian_opt/HartreeExchangeGradient.h:44-49), the trace identity E = 1/2 sum_q [n_q h_qq + F_qq],
  and central differences of the energy along orbital rotations U = exp(-kappa) (the orbital
  gradient g_pq = 2 (F_qp - conj(F_pq)) of src/Hessian_opt/OrbitalGradient.h).
- toolkit / pairContract / diagContract / accumU: numpy restatements of their definitions.
- sigma-vector products: the symmetric part of the finite-difference Jacobian of the orbital
  gradient (the true second derivative of the energy). This pins the JK-only products for any
  couplings and the Hartree/exchange products when two_rdm_h == two_rdm_x without L1/L2 terms; for
  two_rdm_h != two_rdm_x and for the L1/L2 terms the reference's RI form itself departs from the
  finite-difference Hessian, so there the oracle is a line-by-line restatement (parity unpinned).
"""
import numpy as np
import pytest

from oracle import oracle  # noqa: E402

import _dense as dn  # noqa: E402


def rnd_case(seed, n, naux, cplx):
    g = np.random.default_rng(seed)
    b = g.standard_normal((naux, n, n))
    if cplx:
        b = b + 1j * g.standard_normal((naux, n, n))
    b = 0.5 * (b + np.conj(np.transpose(b, (0, 2, 1))))
    h = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0.0)
    h = 0.5 * (h + np.conj(h.T))
    occ = g.uniform(0.1, 0.9, n)
    h2 = g.standard_normal((n, n))
    x2 = g.standard_normal((n, n))
    return g, b, h, occ, 0.5 * (h2 + h2.T), 0.5 * (x2 + x2.T)


def pair_sym(m, pair):
    m = 0.5 * (m + m.T)
    return 0.5 * (m + m[np.ix_(pair, pair)])


PAIR6 = np.array([1, 0, 3, 2, 5, 4])


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("with_l", [False, True])
def test_hx_fock_definition_trace_and_gradient(cplx, with_l):
    g, b, h, occ, h2, x2 = rnd_case(31 + cplx + 2 * with_l, 6, 9, cplx)
    kw = {}
    if with_l:
        kw = dict(pair=PAIR6, l1=pair_sym(g.standard_normal((6, 6)), PAIR6), l2=pair_sym(g.standard_normal((6, 6)), PAIR6))
    f = oracle.ri_hx_fock_gemm(b.reshape(9, -1), h, occ, h2, x2, kw.get("pair"), kw.get("l1"), kw.get("l2"))
    fd = dn.fock(h, b, occ, h2, x2, **kw)
    assert np.max(np.abs(f - fd)) < 1e-12 * max(1.0, np.max(np.abs(fd)))
    e = dn.energy(h, b, occ, h2, x2, **kw)
    assert abs(e - 0.5 * np.real(np.sum(occ * np.diag(h)) + np.trace(f))) < 1e-11 * max(1.0, abs(e))
    pairs = dn.lower_pairs(6)
    v = g.standard_normal(len(pairs) * (2 if cplx else 1))
    kap = dn.kappa_joint(6, pairs, v) if cplx else dn.kappa_real(6, pairs, v)
    de = dn.d_dt(lambda t: dn.energy(*dn.rotate(h, b, t * kap), occ, h2, x2, **kw))
    assert abs(de - dn.gradient_vector(f, pairs, joint=cplx) @ v) < 1e-7 * max(1.0, abs(de))


def test_hx_fock_planted_l_factor_is_caught(monkeypatch):
    g, b, h, occ, h2, x2 = rnd_case(40, 6, 7, False)
    l1, l2 = pair_sym(g.standard_normal((6, 6)), PAIR6), pair_sym(g.standard_normal((6, 6)), PAIR6)
    monkeypatch.setenv("RR_ORACLE_PLANT", "hx_l_factor")
    f = oracle.ri_hx_fock_gemm(b.reshape(7, -1), h, occ, h2, x2, PAIR6, l1, l2)
    fd = dn.fock(h, b, occ, h2, x2, pair=PAIR6, l1=l1, l2=l2)
    assert np.max(np.abs(f - fd)) > 1e-3


@pytest.mark.parametrize("cplx", [False, True])
def test_toolkit_and_primitives_match_definitions(cplx):
    g, b, h, occ, h2, x2 = rnd_case(50 + cplx, 5, 7, cplx)
    n, naux = 5, 7
    k = g.standard_normal((n, n))
    kb1, kb2, bd, k1d, k2d = oracle.ri_sigma_toolkit(b.reshape(naux, -1), k)
    assert np.allclose(kb1, np.einsum("as,Psb->Pab", k, b), atol=1e-13)
    assert np.allclose(kb2, np.einsum("Pas,sb->Pab", b, k), atol=1e-13)
    assert np.allclose(bd, np.einsum("Ptt->Pt", b)) and np.allclose(k1d, np.einsum("Ptt->Pt", kb1))
    assert np.allclose(k2d, np.einsum("Ptt->Pt", kb2))
    # pairContract with index tables, weights, row factors and a post factor on both sides
    idx_f = g.permutation(n * n).reshape(n, n).astype(np.uint32)
    idx_g = g.permutation(n * n).reshape(n, n).astype(np.uint32)
    wf, wg = g.standard_normal((n, n)), g.standard_normal((n, n))
    row = g.standard_normal((naux, n)) + (1j * g.standard_normal((naux, n)) if cplx else 0.0)
    post = g.standard_normal((n, n))
    t0 = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0.0)
    bf = b.reshape(naux, -1)
    F = bf[:, idx_f] * wf[None] * row[:, :, None]          # F(P,t,x)
    F = np.einsum("Pta,ax->Ptx", F, post)
    G = kb1.reshape(naux, -1)[:, idx_g] * wg[None]
    want = t0 - 0.7 * np.einsum("Ptx,Pty->xy", F, G)
    got = oracle.ri_pair_contract(t0.copy(), -0.7, n, naux, dict(s=bf, idx=idx_f, w=wf, row=row, post=post),
                                  dict(s=kb1.reshape(naux, -1), idx=idx_g, w=wg))
    assert np.allclose(got, want, atol=1e-12)
    ua = g.standard_normal((naux, n)) + (1j * g.standard_normal((naux, n)) if cplx else 0.0)
    ub = g.standard_normal((naux, n))
    want = t0 + np.einsum("Pba,Pab->ab", b, ua[:, :, None] + ub[:, None, :])
    assert np.allclose(oracle.ri_diag_contract(t0.copy(), b, ua, ub), want, atol=1e-12)
    wm = g.standard_normal((n, n))
    assert np.allclose(oracle.ri_accum_u(ua.copy(), 0.3, row, wm), ua + 0.3 * row @ wm.T, atol=1e-12)


def fd_hessian(h, b, occ, h2, x2, pairs, joint, **kw):
    m = len(pairs) * (2 if joint else 1)
    n = h.shape[0]
    H = np.zeros((m, m))
    for j in range(m):
        e = np.zeros(m)
        e[j] = 1.0
        kap = dn.kappa_joint(n, pairs, e) if joint else dn.kappa_real(n, pairs, e)
        H[:, j] = dn.d_dt(lambda t: dn.gradient_vector(dn.fock(*dn.rotate(h, b, t * kap), occ, h2, x2, **kw), pairs,
                                                        joint=joint))
    return 0.5 * (H + H.T)


@pytest.mark.parametrize("cplx", [False, True])
def test_jk_sigma_is_the_energy_hessian(cplx):
    g, b, h, occ, h2, x2 = rnd_case(60 + cplx, 5, 6, cplx)
    pairs = dn.lower_pairs(5)
    hs = fd_hessian(h, b, occ, h2, x2, pairs, cplx)
    v = g.standard_normal(hs.shape[0])
    w = oracle.sigma_hessian_vector_ri(h, b, occ, h2, x2, pairs, v)
    assert np.max(np.abs(w - hs @ v)) < 1e-6 * max(1.0, np.max(np.abs(hs @ v)))


@pytest.mark.parametrize("cplx", [False, True])
def test_hx_sigma_is_the_energy_hessian_for_equal_couplings(cplx):
    g, b, h, occ, h2, _ = rnd_case(70 + cplx, 5, 6, cplx)
    pairs = dn.lower_pairs(5)
    hs = fd_hessian(h, b, occ, h2, h2, pairs, cplx)
    f = dn.fock(h, b, occ, h2, h2)
    v = g.standard_normal(hs.shape[0])
    w = oracle.sigma_hessian_vector_ri(h, b, occ, h2, h2, pairs, v, fock=f, hx=True)
    assert np.max(np.abs(w - hs @ v)) < 1e-6 * max(1.0, np.max(np.abs(hs @ v)))
    # and equals the JK-only product for the same (equal) couplings
    w_jk = oracle.sigma_hessian_vector_ri(h, b, occ, h2, h2, pairs, v)
    assert np.max(np.abs(w - w_jk)) < 1e-10 * max(1.0, np.max(np.abs(w)))


@pytest.mark.parametrize("plant", ["sigma_post", "sigma_row"])
def test_sigma_planted_mistakes_are_caught(monkeypatch, plant):
    g, b, h, occ, h2, x2 = rnd_case(80, 5, 6, False)
    pairs = dn.lower_pairs(5)
    hs = fd_hessian(h, b, occ, h2, x2, pairs, False)
    v = g.standard_normal(hs.shape[0])
    monkeypatch.setenv("RR_ORACLE_PLANT", plant)
    w = oracle.sigma_hessian_vector_ri(h, b, occ, h2, x2, pairs, v)
    assert np.max(np.abs(w - hs @ v)) > 1e-3
