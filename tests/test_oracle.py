"""Pins the CPU oracle (oracle/rr_oracle.cpp) before it is trusted as the GPU checker.

Every oracle function is checked against a formula it does not itself use: the dense 4-index
tensor (pq|rs) = sum_P B_P(p,q) B_P(r,s) contracted element by element (the reference's own test
pattern, tests/test_ao_cholesky.cpp:50-59, :61-83, :85-120), four-leg MO transforms of that dense
tensor, closed-form invariants (recovery eigenvalues, rotation round trips, Hermiticity), and
planted mistakes that each check must catch.
"""
import numpy as np
import pytest

from rerdmft_amd import synthetic as syn
from tests._lcg import Lcg


def dense_eri(b):
    """(pq|rs) from a (naux, n, n) tensor."""
    return np.einsum("Ppq,Prs->pqrs", b, b)


def brute_nonrel_fock(h, eri, p):
    # F(a,r) = h + sum_qs P(s,q) (ar|qs) - 1/2 sum_qs P(s,q) (as|qr)   (test_ao_cholesky.cpp:50-59)
    n = h.shape[0]
    f = h.copy()
    for a in range(n):
        for r in range(n):
            hart = 0.0
            exch = 0.0
            for q in range(n):
                for s in range(n):
                    hart += p[s, q] * eri[a, r, q, s]
                    exch += p[s, q] * eri[a, s, q, r]
            f[a, r] += hart - 0.5 * exch
    return f


def spin_eri(eri, n):
    """Closed-shell spin-orbital <AB|CD> = (AC|BD) if spin(A)=spin(C), spin(B)=spin(D)
    (test_ao_cholesky.cpp:72-76), as a dense (2n)^4 array."""
    n2 = 2 * n
    out = np.zeros((n2, n2, n2, n2))
    for sa in range(2):
        for sb in range(2):
            out[sa * n:(sa + 1) * n, sb * n:(sb + 1) * n, sa * n:(sa + 1) * n, sb * n:(sb + 1) * n] = \
                np.transpose(eri, (0, 2, 1, 3))
    return out


def brute_x2c_fock(h, seri, p):
    # F(a,r) = h + sum P(s,q) <aq|rs> - sum P(s,q) <aq|sr>   (test_ao_cholesky.cpp:77-82)
    n2 = h.shape[0]
    f = h.astype(np.complex128).copy()
    for a in range(n2):
        for r in range(n2):
            f[a, r] += np.sum(p.T * seri[a, :, r, :]) - np.sum(p.T * seri[a, :, :, r])
    return f


def make_lcg_ao(n, rank, rng):
    """Rank-`rank` symmetric vectors drawn exactly as tests/test_ao_cholesky.cpp:29-32."""
    b = np.zeros((rank, n, n))
    for m in range(rank):
        for i in range(n):
            for j in range(i, n):
                b[m, i, j] = b[m, j, i] = 0.3 * rng.next()
    return b


# ---------------------------------------------------------------------------------------------
# Cholesky Fock on the reference's own seeded inputs (test_ao_cholesky.cpp:28-83)
# ---------------------------------------------------------------------------------------------
def lcg_ao_inputs():
    rng = Lcg(20260926)
    n, rank = 6, 9
    vec = make_lcg_ao(n, rank, rng)
    h = np.zeros((n, n))
    p = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            h[i, j] = h[j, i] = rng.next()
            p[i, j] = p[j, i] = rng.next()
    n2 = 2 * n
    hx = np.zeros((n2, n2), dtype=np.complex128)
    px = np.zeros((n2, n2), dtype=np.complex128)
    for i in range(n2):
        for j in range(i, n2):
            v = complex(rng.next(), 0.0 if i == j else rng.next())
            hx[i, j] = v
            hx[j, i] = np.conj(v)
            w = complex(rng.next(), 0.0 if i == j else rng.next())
            px[i, j] = w
            px[j, i] = np.conj(w)
    return vec, h, p, hx, px, rng


def test_cholesky_nonrel_fock_matches_dense(oracle):
    vec, h, p, _, _, _ = lcg_ao_inputs()
    f = oracle.nonrel_fock_cholesky(vec.reshape(len(vec), -1), h, p)
    ref = brute_nonrel_fock(h, dense_eri(vec), p)
    assert np.max(np.abs(f - ref)) < 1e-9


def test_cholesky_x2c_fock_matches_dense(oracle):
    vec, _, _, hx, px, _ = lcg_ao_inputs()
    n = vec.shape[1]
    f = oracle.x2c_fock_cholesky(vec.reshape(len(vec), -1), hx, px)
    ref = brute_x2c_fock(hx, spin_eri(dense_eri(vec), n), px)
    assert np.max(np.abs(f - ref)) < 1e-9


def test_cholesky_mo_vectors_match_dense_transform(oracle):
    vec, _, _, _, _, rng = lcg_ao_inputs()
    n, nmo = vec.shape[1], 5
    cr = np.array([[rng.next() for _ in range(nmo)] for _ in range(n)])
    w = oracle.cholesky_to_mo_spatial(vec.reshape(len(vec), -1), cr)
    mo = np.einsum("Lpq,Lrs->pqrs", w, w)  # (pa pc|pb pd) from vectors
    ref = np.einsum("mnls,ma,nc,lb,sd->acbd", dense_eri(vec), cr, cr, cr, cr)
    assert np.max(np.abs(mo - ref)) < 1e-9
    cc = np.array([[complex(rng.next(), rng.next()) for _ in range(nmo)] for _ in range(2 * n)])
    ws = oracle.cholesky_to_mo_spinor(vec.reshape(len(vec), -1), cc)
    # <ab|cd> = sum_L W_L(c... ) per CholeskyEri convention: eri(a,b,c,d) = sum_L V_L(a,c) V_L(b,d), V = W^T
    v = np.transpose(ws, (0, 2, 1))
    got = np.einsum("Lac,Lbd->abcd", v, v)
    seri = spin_eri(dense_eri(vec), n)
    ref = np.einsum("ABCD,Aa,Bb,Cc,Dd->abcd", seri, cc.conj(), cc.conj(), cc, cc)
    assert np.max(np.abs(got - ref)) < 1e-9


# ---------------------------------------------------------------------------------------------
# RI NON_REL with occupied-orbital recovery (NonRelFockMatrixRi.cpp:27-100)
# ---------------------------------------------------------------------------------------------
def nonrel_case(n=8, naux=13, nocc=3, seed=7):
    b = syn.ri_tensor(naux, n, 0.3, 4.0, seed)
    s = syn.spd_overlap(n, seed + 1)
    x = syn.inverse_sqrt(s)
    c = syn.s_orthonormal_occ(s, nocc, seed + 2)
    p = syn.density(c)
    h = syn.symmetric(n, seed + 3)
    return b, s, x, c, p, h


def test_nonrel_ri_fock_matches_dense_for_idempotent_density(oracle):
    b, s, x, c, p, h = nonrel_case()
    f, j, k = oracle.nonrel_fock_ri(b.reshape(len(b), -1), x, s, h, p, return_jk=True)
    ref = brute_nonrel_fock(h, dense_eri(b), p)
    assert np.max(np.abs(f - ref)) < 1e-11
    # J is linear in P and K = sum_P B P B for idempotent P
    assert np.max(np.abs(j - np.einsum("Pmn,Pqs,qs->mn", b, b, p))) < 1e-11
    assert np.max(np.abs(k - k.T)) < 1e-12


def test_nonrel_ri_recovery_keeps_only_eigenvalues_above_half(oracle):
    # Non-idempotent P: eigenvalues of X^T S P S X / 2 at 0.9 and 0.3 -> only the 0.9 direction is kept
    # and it is weighted as doubly occupied (sqrt(2) folding, NonRelFockMatrixRi.cpp:67-69).
    b, s, x, c, _, h = nonrel_case(nocc=2)
    p = 2.0 * (0.9 * np.outer(c[:, 0], c[:, 0]) + 0.3 * np.outer(c[:, 1], c[:, 1]))
    _, _, k = oracle.nonrel_fock_ri(b.reshape(len(b), -1), x, s, h, p, return_jk=True)
    p_kept = 2.0 * np.outer(c[:, 0], c[:, 0])
    k_ref = np.einsum("Pam,mn,Pnb->ab", b, p_kept, b)
    assert np.max(np.abs(k - k_ref)) < 1e-11


def test_factor_form_matches_bdb_for_any_factor(oracle):
    b, _, _, c, p, _ = nonrel_case()
    g = syn.rng(3)
    l, r = g.standard_normal((8, 4)), g.standard_normal((8, 4))
    j, k = oracle.ri_jk_factor(b.reshape(len(b), -1), p, l, r)
    k_ref = np.einsum("Pma,mn,Pnb->ab", b, l @ r.T, b)
    assert np.max(np.abs(k - k_ref)) < 1e-11
    assert np.max(np.abs(j - np.einsum("Pmn,Pqs,qs->mn", b, b, p))) < 1e-11


# ---------------------------------------------------------------------------------------------
# X2C RI (X2C_FockMatrixRi.cpp:35-162)
# ---------------------------------------------------------------------------------------------
def x2c_case(nl=5, naux=11, npairs=2, seed=11):
    b = syn.ri_tensor(naux, nl, 0.3, 3.0, seed)
    s1 = syn.spd_overlap(nl, seed + 1)
    s = np.kron(np.eye(2), s1).astype(np.complex128)
    x = np.kron(np.eye(2), syn.inverse_sqrt(s1)).astype(np.complex128)
    q = syn.kramers_pair_occ(nl, npairs, seed + 2)
    c = x @ q  # C^H S C = I
    p = c @ c.conj().T
    h = syn.hermitian(2 * nl, seed + 3)
    return b, s, x, c, p, h


def test_x2c_ri_fock_matches_spin_orbital_dense(oracle):
    b, s, x, c, p, h = x2c_case()
    nl = b.shape[1]
    f = oracle.x2c_fock_ri(b.reshape(len(b), -1), x, s, h, p)
    ref = brute_x2c_fock(h, spin_eri(dense_eri(b), nl), p)
    assert np.max(np.abs(f - ref)) < 1e-11
    assert np.max(np.abs(f - f.conj().T)) < 1e-12


# ---------------------------------------------------------------------------------------------
# 4c RKB (UkbFockMatrixRi.cpp:35-216)
# ---------------------------------------------------------------------------------------------
def ukb_dense(bll, bss, nl, ns):
    """Dense UKB <AB|CD> = (AC|BD): A,C in one flavor, B,D in one flavor; spatial integrals from the
    type (L/S) of each flavor."""
    offs = [0, nl, 2 * nl, 2 * nl + ns]
    sizes = [nl, nl, ns, ns]
    types = [bll, bll, bss, bss]
    n = 2 * nl + 2 * ns
    out = np.zeros((n, n, n, n))
    for f in range(4):
        for u in range(4):
            blk = np.einsum("Pac,Pbd->abcd", types[f], types[u])
            out[offs[f]:offs[f] + sizes[f], offs[u]:offs[u] + sizes[u],
                offs[f]:offs[f] + sizes[f], offs[u]:offs[u] + sizes[u]] = blk
    return out


def rkb_case(nl=3, ns=4, naux=7, nocc=3, seed=21):
    bll = syn.ri_tensor(naux, nl, 0.3, 3.0, seed)
    bss = syn.ri_tensor(naux, ns, 0.2, 3.0, seed + 1)
    n = 2 * nl + 2 * ns
    g = syn.rng(seed + 2)
    c = g.standard_normal((n, nocc)) + 1j * g.standard_normal((n, nocc))
    occ = 0.6 + 0.4 * g.random(nocc)
    dens = syn.hermitian(n, seed + 3, 0.3)
    return bll, bss, c, occ, dens


def test_ri_fock_two_electron_direct_matches_dense(oracle):
    nl, ns = 3, 4
    bll, bss, c, occ, dens = rkb_case(nl, ns)
    g = oracle.ri_fock_two_electron_direct(nl, ns, bll.reshape(len(bll), -1), bss.reshape(len(bss), -1), dens, c, occ)
    eri = ukb_dense(bll, bss, nl, ns)
    d_occ = (c * occ[None, :]) @ c.conj().T
    n = 2 * nl + 2 * ns
    ref = np.zeros((n, n), dtype=np.complex128)
    # J(A,R) = sum P(S,Q) <AQ|RS> ; K(A,R) = sum D(S,Q) <AQ|SR>
    for a in range(n):
        for r in range(n):
            ref[a, r] = np.sum(dens.real.T * eri[a, :, r, :]) - np.sum(d_occ.T * eri[a, :, :, r])
    assert np.max(np.abs(g - ref)) < 1e-11


def test_rkb_fock_recovery_and_embedding(oracle):
    # rkbFockMatrix with v_total = I (n_rkb = n_ukb) and an idempotent RKB density reduces to
    # h + riFockTwoElectronDirect(P, C_occ, 1).
    nl, ns, naux, nocc = 3, 4, 7, 3
    bll, bss, _, _, _ = rkb_case(nl, ns, naux, nocc)
    n = 2 * nl + 2 * ns
    sr = syn.spd_overlap(n, 5).astype(np.complex128)
    x = syn.inverse_sqrt(sr.real).astype(np.complex128)
    q = syn.unitary(n, 6)[:, :nocc]
    c = x @ q
    p = c @ c.conj().T
    h = syn.hermitian(n, 7)
    v = np.eye(n, dtype=np.complex128)
    f = oracle.rkb_fock_ri(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), v, sr, x, h, p)
    g = oracle.ri_fock_two_electron_direct(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), p, c, np.ones(nocc))
    assert np.max(np.abs(f - (h + g))) < 1e-10


def rkb_rect_case(nl=3, ns=4, naux=7, n_rkb=10, nocc=3, seed=41):
    """A rectangular complex RKB→UKB map V (n_ukb × n_rkb, n_ukb = 2nl + 2ns = 14) and an idempotent
    RKB density P = C C^H with C^H S C = I."""
    bll, bss, _, _, _ = rkb_case(nl, ns, naux, nocc, seed)
    n_ukb = 2 * nl + 2 * ns
    g = syn.rng(seed + 10)
    v = g.standard_normal((n_ukb, n_rkb)) + 1j * g.standard_normal((n_ukb, n_rkb))
    a = g.standard_normal((n_rkb, n_rkb)) + 1j * g.standard_normal((n_rkb, n_rkb))
    s = a @ a.conj().T / n_rkb + np.eye(n_rkb)
    w, u = np.linalg.eigh(s)
    x = (u / np.sqrt(w)) @ u.conj().T
    c = x @ syn.unitary(n_rkb, seed + 11)[:, :nocc]
    h = syn.hermitian(n_rkb, seed + 12)
    return bll, bss, v, s, x, c, c @ c.conj().T, h


def dense_ukb_g(bll, bss, nl, ns, p_ukb, d_occ):
    """G(A,R) = sum P(S,Q) <AQ|RS> - sum D(S,Q) <AQ|SR> on the dense four-flavor tensor."""
    eri = ukb_dense(bll, bss, nl, ns)
    return np.einsum("sq,aqrs->ar", p_ukb.real, eri) - np.einsum("sq,aqsr->ar", d_occ, eri)


def test_rkb_fock_rectangular_v_matches_dense_embedding(oracle):
    # F = h + V^H G[V P V^H] V with G from the dense UKB tensor and the occupied set V C (an
    # idempotent RKB density recovers C up to a unitary, which K does not see).
    nl, ns, naux = 3, 4, 7
    bll, bss, v, s, x, c, p, h = rkb_rect_case(nl, ns, naux)
    f = oracle.rkb_fock_ri(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), v, s, x, h, p)
    p_ukb = v @ p @ v.conj().T
    g = dense_ukb_g(bll, bss, nl, ns, p_ukb, (v @ c) @ (v @ c).conj().T)
    ref = h + v.conj().T @ g @ v
    assert np.max(np.abs(f - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(f - f.conj().T)) < 1e-10 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("plant", ["rkb_vt", "rkb_pvt"])
def test_rkb_fock_planted_dagger_is_rejected(oracle, monkeypatch, plant):
    nl, ns, naux = 3, 4, 7
    bll, bss, v, s, x, c, p, h = rkb_rect_case(nl, ns, naux)
    p_ukb = v @ p @ v.conj().T
    ref = h + v.conj().T @ dense_ukb_g(bll, bss, nl, ns, p_ukb, (v @ c) @ (v @ c).conj().T) @ v
    monkeypatch.setenv("RR_ORACLE_PLANT", plant)
    f = oracle.rkb_fock_ri(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), v, s, x, h, p)
    assert np.max(np.abs(f - ref)) > 1e-3 * max(1.0, np.max(np.abs(ref)))


# ---------------------------------------------------------------------------------------------
# MO-basis tensors (RiMoEri.cpp:20-121, NonRelSpinRiMoEri.cpp:13-95)
# ---------------------------------------------------------------------------------------------
def test_build_ri_mo_eri_matches_four_leg_transform(oracle):
    nl, ns, naux = 3, 2, 6
    bll = syn.ri_tensor(naux, nl, 0.3, 3.0, 31)
    bss = syn.ri_tensor(naux, ns, 0.3, 3.0, 32)
    n = 2 * nl + 2 * ns
    g = syn.rng(33)
    nt = 5
    c = g.standard_normal((n, nt)) + 1j * g.standard_normal((n, nt))
    bmo = oracle.build_ri_mo_eri(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), c)
    got = np.einsum("Pac,Pbd->abcd", bmo, bmo)  # eri(a,b,c,d) = sum_P B(P,a,c) B(P,b,d)
    ref = np.einsum("ABCD,Aa,Bb,Cc,Dd->abcd", ukb_dense(bll, bss, nl, ns), c.conj(), c.conj(), c, c)
    assert np.max(np.abs(got - ref)) < 1e-11


def test_ri_mo_eri_rotation_round_trip_and_transform(oracle):
    g = syn.rng(41)
    naux, nt = 5, 4
    b = g.standard_normal((naux, nt, nt)) + 1j * g.standard_normal((naux, nt, nt))
    u = syn.unitary(nt, 42)
    br = oracle.ri_mo_eri_rotated(b, u)
    back = oracle.ri_mo_eri_rotated(br, u.conj().T)
    assert np.max(np.abs(back - b)) < 1e-12
    ref = np.einsum("ap,Pab,bq->Ppq", u.conj(), b, u)
    assert np.max(np.abs(br - ref)) < 1e-12


def test_nonrel_spin_mo_eri_and_rotation(oracle):
    naux, n = 6, 5
    b = syn.ri_tensor(naux, n, 0.3, 3.0, 51)
    c = syn.orthonormal_occ(n, n, 52)
    bmo = oracle.build_ri_nonrel_spin_mo_eri(b.reshape(naux, -1), c)
    ref = np.einsum("mnls,ma,nc,lb,sd->acbd", dense_eri(b), c, c, c, c)
    assert np.max(np.abs(np.einsum("Pac,Pbd->acbd", bmo, bmo) - ref)) < 1e-11
    us = syn.orthonormal_occ(n, n, 53)
    u = np.kron(np.eye(2), us)
    rot = oracle.ri_nonrel_spin_mo_eri_rotated(bmo, u)
    assert np.max(np.abs(rot - np.einsum("ap,Pab,bq->Ppq", us, bmo, us))) < 1e-12
    bad = u.copy()
    bad[0, n] = 0.1
    with pytest.raises(RuntimeError, match="not spin-restricted"):
        oracle.ri_nonrel_spin_mo_eri_rotated(bmo, bad)


# ---------------------------------------------------------------------------------------------
# Planted mistakes: each named mistake is switched on inside the oracle (RR_ORACLE_PLANT) and the
# pinning check for that function must then fail.
# ---------------------------------------------------------------------------------------------
PLANTS = {
    "half": test_nonrel_ri_fock_matches_dense_for_idempotent_density,
    "sqrt2": test_nonrel_ri_fock_matches_dense_for_idempotent_density,
    "threshold": test_nonrel_ri_recovery_keeps_only_eigenvalues_above_half,
    "conj": test_x2c_ri_fock_matches_spin_orbital_dense,
    "mo_conj": test_build_ri_mo_eri_matches_four_leg_transform,
    "rot_dagger": test_ri_mo_eri_rotation_round_trip_and_transform,
}


@pytest.mark.parametrize("name", sorted(PLANTS))
def test_planted_mistake_is_caught(oracle, monkeypatch, name):
    monkeypatch.setenv("RR_ORACLE_PLANT", name)
    with pytest.raises(AssertionError):
        PLANTS[name](oracle)


def test_planted_transpose_caught_by_cholesky_check(oracle, monkeypatch):
    # P B instead of P B^T is invisible for symmetric B; the LCG vectors are symmetric, so use a
    # non-symmetric P to expose it against the dense formula.
    monkeypatch.setenv("RR_ORACLE_PLANT", "chol_transpose")
    vec, h, p, _, _, _ = lcg_ao_inputs()
    vec = vec.copy()
    vec[:, 0, 1] += 0.1  # break symmetry of the vectors
    f = oracle.nonrel_fock_cholesky(vec.reshape(len(vec), -1), h, p)
    n = h.shape[0]
    j = np.einsum("Lab,Lqs,sq->ab", vec, vec, p)
    k = np.einsum("Lam,mn,Lnb->ab", vec, p, vec)
    assert np.max(np.abs(f - (h + j - 0.5 * k))) > 1e-6


def test_factor_form_with_recovery_equals_cholesky_density_form(oracle):
    """For an SCF density (P = C occ C^H with C^H S C = I) the RI builds' recovered-factor exchange
    equals the density-form exchange sum_L B_L P B_L of AoCholesky.cpp:108-217 on the same vectors:
    the identity behind running the factor-form builds on Cholesky slots."""
    from tests.test_gpu_relativistic import x2c_inputs
    nl, naux = 20, 60
    b, s, x, p, h = x2c_inputs(nl, naux, 4, 401)
    f_ri = oracle.x2c_fock_ri(b.reshape(naux, -1), x, s, h, p)
    f_ch = oracle.x2c_fock_cholesky(b.reshape(naux, -1), h, p)
    assert np.max(np.abs(f_ri - f_ch)) < 1e-12 * np.max(np.abs(f_ch))
    n = 30
    bb = syn.ri_tensor(naux, n, 0.5, 8.0, 402)
    s1 = syn.spd_overlap(n, 403)
    xx = syn.inverse_sqrt(s1)
    c = xx @ syn.orthonormal_occ(n, 5, 404)
    pp, hh = 2 * c @ c.T, syn.symmetric(n, 405)
    g_ri = oracle.nonrel_fock_ri(bb.reshape(naux, -1), xx, s1, hh, pp)
    g_ch = oracle.nonrel_fock_cholesky(bb.reshape(naux, -1), hh, pp)
    assert np.max(np.abs(g_ri - g_ch)) < 1e-12 * np.max(np.abs(g_ch))
    # a density that is not idempotent in the S metric (occupation 0.7) is outside the identity
    p2 = 0.7 * c @ c.T
    assert np.max(np.abs(oracle.nonrel_fock_ri(bb.reshape(naux, -1), xx, s1, hh, p2)
                         - oracle.nonrel_fock_cholesky(bb.reshape(naux, -1), hh, p2))) > 1e-6
