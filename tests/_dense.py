"""Dense (four-index) restatements used to pin the functional-stage oracle at tiny sizes.

eri(a,b,c,d) = sum_P B(P,a,c) B(P,b,d) (physicist notation, unconjugated), energy and generalized
Fock of the Hartree/exchange(+L1/L2) ansatz as defined in the reference's documentation
(src/Hessian_opt/HartreeExchangeGradient.h:44-49 and :90-96), orbital rotations U = exp(-kappa)
(src/Hessian_opt/SpinorRotation.h) applied as B'_P = U^H B_P U, h' = U^H h U.
"""
import numpy as np
import scipy.linalg


def eri(b):
    return np.einsum("Pac,Pbd->abcd", b, b)


def energy(h, b, occ, h2, x2, pair=None, l1=None, l2=None):
    g = eri(b)
    n = h.shape[0]
    idx = np.arange(n)
    e = np.sum(occ * np.diag(h))
    e += 0.5 * np.sum(g[idx[:, None], idx[None, :], idx[:, None], idx[None, :]] * h2)
    e -= 0.5 * np.sum(g[idx[:, None], idx[None, :], idx[None, :], idx[:, None]] * x2)
    if pair is not None:
        pr = np.asarray(pair)
        p, q = idx[:, None], idx[None, :]
        e += np.sum(g[p, pr[p], q, pr[q]] * l1) + np.sum(g[p, pr[p], pr[q], q] * l2)
    return float(np.real(e))


def fock(h, b, occ, h2, x2, pair=None, l1=None, l2=None):
    """F_pq = occ_q h(q,p) + sum_s [h2(s,q) <qs|ps> - x2(q,s) <qs|sp>] + 2 sum_s <s sbar|qbar p>(l1(s,qbar)+l2(s,q))."""
    g = eri(b)
    n = h.shape[0]
    f = occ[None, :] * h.T
    f = f + np.einsum("sq,qsps->pq", h2, g) - np.einsum("qs,qssp->pq", x2, g)
    if pair is not None:
        pr = np.asarray(pair)
        s = np.arange(n)
        c = l1[:, pr] + l2                       # c(s,q) = l1(s,qbar) + l2(s,q)
        gs = g[s, pr[s]]                         # gs[s, c', d] = <s sbar|c' d>
        t = gs[:, pr, :]                         # t[s, q, p] = <s sbar|qbar p>
        f = f + 2.0 * np.einsum("sqp,sq->pq", t, c)
    return f


def rotate(h, b, kappa):
    u = scipy.linalg.expm(-kappa)
    hr = u.conj().T @ h @ u
    br = np.einsum("am,Pmn,nc->Pac", u.conj().T, b, u)
    return hr, br


def kappa_real(n, pairs, v):
    k = np.zeros((n, n))
    for (p, q), x in zip(pairs, v):
        k[p, q] = x
        k[q, p] = -x
    return k


def kappa_joint(n, pairs, v):
    m = len(pairs)
    k = np.zeros((n, n), dtype=complex)
    for i, (p, q) in enumerate(pairs):
        k[p, q] += v[i] + 1j * v[m + i]
        k[q, p] += -v[i] + 1j * v[m + i]
    return k


def gradient_vector(f, pairs, joint=False):
    """g_pq = 2 (F_qp - conj(F_pq)) (src/Hessian_opt/OrbitalGradient.h); real part (and imaginary for joint)."""
    g = np.array([2.0 * (f[q, p] - np.conj(f[p, q])) for p, q in pairs])
    return np.concatenate([g.real, g.imag]) if joint else g.real


def d_dt(fun, step=1e-3):
    """Fourth-order central difference of fun(t) at 0."""
    return (8.0 * (fun(step) - fun(-step)) - (fun(2 * step) - fun(-2 * step))) / (12.0 * step)


def lower_pairs(n):
    return [(p, q) for p in range(n) for q in range(p)]
