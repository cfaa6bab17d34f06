"""FP64 CPU restatement of the reference's CASSCF RI contractions (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of rerdmft_amd/csrc/rr_cas.hip; never by
the library.

Follows src/MCSCF/CasFock.cpp (RI form, <ab|cd> = sum_P B(P, a n + c) B(P, b n + d), no conjugation):
  two_electron_operator  riTwoElectronOperator (:73-113): J_rp = sum_P c_P B(P, r n + p),
                         c_P = sum_st rho_st B(P, s n + t); rho = U diag(w) U^H with |w| > 1e-14 kept
                         (lowRankFactors, :45-68); K(r,p) = sum_P [B_P conj(U)](r,k) [diag(w) U^T B_P](k,p);
                         G = J - K.
  active_integrals       riActiveIntegrals (:136-148): Ba^T Ba, Ba[P][(t,u)] = B(P, act_t n + act_u).
  active_cumulant_fock   riActiveCumulantFock (:115-134).

Pins (tests/test_oracle_cas.py): all three against the dense four-index tensor built from the same
B (the generic casTwoElectronOperator of CasFock.h:44-58, G_rp = sum_ab rho_ab (<ar|bp> - <ra|bp>),
and the (tu|vw) / cumulant definitions), real and complex, with a planted K sign and a planted
index swap in the cumulant Fock shown to fail.
"""
import numpy as np


def _low_rank(rho):
    w, u = np.linalg.eigh(rho)
    keep = np.abs(w) > 1e-14
    return w[keep], u[:, keep]


def two_electron_operator(b, n, rho, planted=False):
    b = b.reshape(b.shape[0], n * n)
    c = b @ rho.reshape(-1)
    j = (b.T @ c).reshape(n, n)
    w, u = _low_rank(rho)
    k = np.zeros((n, n), dtype=np.result_type(b, rho))
    for bp in b.reshape(-1, n, n):
        k += (bp @ u.conj()) @ (w[:, None] * (u.T @ bp))
    return j + k if planted else j - k


def active_integrals(b, n, act):
    b = b.reshape(b.shape[0], n, n)
    ba = b[:, act][:, :, act].reshape(b.shape[0], -1)
    return ba.T @ ba


def active_cumulant_fock(b, n, act, lam, planted=False):
    b = b.reshape(b.shape[0], n, n)
    a = len(act)
    bact = b[:, act][:, :, act].reshape(b.shape[0], a * a)
    lam2 = np.asarray(lam).reshape(a * a, a * a)
    if planted:  # deliberate mistake for the pinning test: lambda's (r,q) pair transposed
        lam2 = np.asarray(lam).reshape(a * a, a, a).transpose(0, 2, 1).reshape(a * a, a * a)
    tm = bact @ lam2
    f = np.zeros((n, a), dtype=np.result_type(b, lam2))
    for r in range(a):
        f += b[:, act[r], :].T @ tm[:, r * a:(r + 1) * a]
    return f


def physics_block(b, n, p, q, r, s):
    """FactoredCcIntegrals::physicsBlock with the dense pair factor (src/CC/Uccsd.cpp:128-153,
    DensePairFactor :27-45): A[P][(p,r)], B'[P][(q,s)] gathered from B, x = A^T B', out[p][q][r][s]."""
    b = b.reshape(b.shape[0], n, n)
    a = b[:, p][:, :, r].reshape(b.shape[0], -1)
    bk = b[:, q][:, :, s].reshape(b.shape[0], -1)
    x = (a.T @ bk).reshape(len(p), len(r), len(q), len(s))
    return x.transpose(0, 2, 1, 3)
