"""Gaussian integrals over Cartesian shells of any angular momentum, and the automatic RI auxiliary
basis: the host FP64 inputs of the RI Fock build.

The reference evaluates its AO and three-centre integrals with libcint (cint1e / cint3c2e_cart /
cint2c2e_cart, src/AO_ints/ThreeCenterIntegrals.cpp:18-21, :112-189), which this build does not
have. This module computes the same integrals by the McMurchie-Davidson scheme (Helgaker, Jorgensen
and Olsen, "Molecular Electronic-Structure Theory", chapter 9): a product of two Cartesian Gaussians
is expanded in Hermite Gaussians with the coefficients E^{ij}_t (eq. 9.5.6-9.5.7), overlap and
kinetic integrals follow from E_0 (9.3.9, 9.3.40), and every Coulomb-type integral from the Hermite
integrals R_{tuv} (9.9.18-9.9.20) of the Boys function F_n (9.8.1):
  nuclear attraction   V = -Z 2 pi / p sum_tuv E_t E_u E_v R_tuv(p, P - C)
  (ab|c), (c|d), (ab|cd)  2 pi^(5/2) / (p q sqrt(p + q)) sum E^ab_tuv (-1)^(tau+nu+phi) E^cd_tau,nu,phi R_(t+tau)(u+nu)(v+phi)
with the auxiliary functions of (ab|c) and (c|d) as the limit of a pair with a zero-exponent s partner.

Basis functions are Cartesian components x^i y^j z^k of a shell, ordered as libcint orders them
(i descending, then j descending), each contracted over the shell's primitives with the given
coefficients of normalized primitives and then normalized to unit self-overlap, as the reference's
normalizeCartesianBasis does. The automatic auxiliary basis is the PAUTO / AUTO recipe of
src/AO_basis/AuxiliaryBasis.cpp:40-190 (Yang, Rendell and Frisch, J. Chem. Phys. 127, 074102
(2007); MOLGW's defaults f_sam = 1.5, lmax_inc = 1).
"""
import math
import re

import numpy as np
from scipy.special import gammainc, gammaln

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903
ELEMENTS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10, "Na": 11,
            "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18}
_LETTER_L = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4}


def cart_components(l):
    """(i, j, k) with i + j + k = l, i descending then j descending (libcint's Cartesian order)."""
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


def _dfact(n):
    """(n)!! for n >= -1."""
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


class Shell:
    """One shell: centre (bohr), angular momentum, primitive exponents and contraction coefficients
    (coefficients of normalized primitives, as a basis-set file gives them)."""

    def __init__(self, centre, l, exps, coefs, element=None):
        self.centre = np.asarray(centre, dtype=np.float64)
        self.l = int(l)
        self.exps = np.asarray(exps, dtype=np.float64)
        self.coefs = np.asarray(coefs, dtype=np.float64)
        self.element = element

    @property
    def size(self):
        return (self.l + 1) * (self.l + 2) // 2


# ---- basis-set files (Gaussian94 format, as the reference's examples/*.gbs) --------------------------
def parse_gbs(text):
    """{element symbol: [(l, exponents, coefficients), ...]}; SP shells become an S and a P shell."""
    out = {}
    elem = None
    lines = [ln.strip() for ln in text.splitlines()]
    i = 0
    while i < len(lines):
        ln = lines[i]
        i += 1
        if not ln or ln.startswith("!") or ln.startswith("#"):
            continue
        if ln == "****":
            elem = None
            continue
        tok = ln.split()
        if elem is None and tok[0] in ELEMENTS and (len(tok) == 1 or tok[1] == "0"):
            elem = tok[0]
            out.setdefault(elem, [])
            continue
        if elem is not None and re.fullmatch(r"[SPDFG]+", tok[0]):
            kind, nprim = tok[0], int(tok[1])
            rows = [[float(x.replace("D", "E").replace("d", "e")) for x in lines[i + k].split()] for k in range(nprim)]
            i += nprim
            exps = np.array([r[0] for r in rows])
            for col, letter in enumerate(kind):
                out[elem].append((_LETTER_L[letter], exps, np.array([r[1 + col] for r in rows])))
    return out


def molecule_shells(atoms, basis):
    """atoms: [(symbol, xyz in bohr)]; basis: parse_gbs result. Shells in atom order, file order."""
    shells = []
    for sym, xyz in atoms:
        for l, e, c in basis[sym]:
            shells.append(Shell(xyz, l, e, c, sym))
    return shells


# ---- primitive-level machinery ------------------------------------------------------------------------
def _prim_norm(a, l):
    """Normalization of the primitive x^l exp(-a r^2) (axis-aligned component of a shell)."""
    return (2.0 * a / np.pi) ** 0.75 * (4.0 * a) ** (0.5 * l) / math.sqrt(_dfact(2 * l - 1))


def hermite_e(la, lb, a, b, xab):
    """E[i, j, t, ...] for i <= la, j <= lb, t <= i + j (HJO 9.5.6-9.5.7); a, b, xab arrays (same shape)
    with xab = A_x - B_x. E[0,0,0] = exp(-mu xab^2)."""
    a, b, xab = np.broadcast_arrays(np.asarray(a, float), np.asarray(b, float), np.asarray(xab, float))
    p = a + b
    mu = a * b / p
    xpa = -b / p * xab
    xpb = a / p * xab
    e = np.zeros((la + 1, lb + 1, la + lb + 2) + a.shape)
    e[0, 0, 0] = np.exp(-mu * xab * xab)
    inv2p = 0.5 / p
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = xpa * e[i - 1, j, t]
                    if t > 0:
                        v = v + inv2p * e[i - 1, j, t - 1]
                    v = v + (t + 1) * e[i - 1, j, t + 1]
                else:
                    v = xpb * e[i, j - 1, t]
                    if t > 0:
                        v = v + inv2p * e[i, j - 1, t - 1]
                    v = v + (t + 1) * e[i, j - 1, t + 1]
                e[i, j, t] = v
    return e[:, :, :la + lb + 1]


def boys(nmax, t):
    """F_n(t) for n = 0..nmax, array (nmax+1, *t.shape): F_n(t) = gamma(n+1/2) P(n+1/2, t) / (2 t^(n+1/2))
    (regularized lower incomplete gamma), with the Taylor series near t = 0."""
    t = np.asarray(t, dtype=np.float64)
    out = np.empty((nmax + 1,) + t.shape)
    small = t < 1e-8
    ts = np.where(small, 1.0, t)
    for n in range(nmax + 1):
        big = np.exp(gammaln(n + 0.5) + np.log(gammainc(n + 0.5, ts)) - (n + 0.5) * np.log(ts)) * 0.5
        ser = 1.0 / (2 * n + 1) - t / (2 * n + 3) + t * t / (2.0 * (2 * n + 5))
        out[n] = np.where(small, ser, big)
    return out


def hermite_r(lmax, alpha, pc):
    """R[t, u, v, ...] = R^0_tuv(alpha, PC) for t + u + v <= lmax (HJO 9.9.18-9.9.20); alpha (...),
    pc (..., 3)."""
    alpha = np.asarray(alpha, dtype=np.float64)
    x, y, z = pc[..., 0], pc[..., 1], pc[..., 2]
    t2 = alpha * (x * x + y * y + z * z)
    f = boys(lmax, t2)
    # r[n][t][u][v]
    r = np.zeros((lmax + 1, lmax + 1, lmax + 1, lmax + 1) + alpha.shape)
    m2a = -2.0 * alpha
    for n in range(lmax + 1):
        r[n, 0, 0, 0] = m2a ** n * f[n]
    for n in range(lmax - 1, -1, -1):
        top = lmax - n
        for t in range(top + 1):
            for u in range(top + 1 - t):
                for v in range(top + 1 - t - u):
                    if t == 0 and u == 0 and v == 0:
                        continue
                    if t > 0:
                        val = x * r[n + 1, t - 1, u, v]
                        if t > 1:
                            val = val + (t - 1) * r[n + 1, t - 2, u, v]
                    elif u > 0:
                        val = y * r[n + 1, t, u - 1, v]
                        if u > 1:
                            val = val + (u - 1) * r[n + 1, t, u - 2, v]
                    else:
                        val = z * r[n + 1, t, u, v - 1]
                        if v > 1:
                            val = val + (v - 1) * r[n + 1, t, u, v - 2]
                    r[n, t, u, v] = val
    return r[0]


# ---- shell-pair data ---------------------------------------------------------------------------------
class _Pair:
    """Primitive-pair data of two shells: p, P, contraction weights and Hermite coefficients per axis."""

    def __init__(self, sa, sb):
        a = sa.exps[:, None] * np.ones((1, len(sb.exps)))
        b = np.ones((len(sa.exps), 1)) * sb.exps[None, :]
        self.la, self.lb = sa.l, sb.l
        self.p = (a + b).ravel()
        self.pc = ((a[..., None] * sa.centre + b[..., None] * sb.centre) / (a + b)[..., None]).reshape(-1, 3)
        w = (sa.coefs * _prim_norm(sa.exps, sa.l))[:, None] * (sb.coefs * _prim_norm(sb.exps, sb.l))[None, :]
        self.w = w.ravel()
        ab = sa.centre - sb.centre
        self.e = [hermite_e(sa.l, sb.l, a.ravel(), b.ravel(), np.full(a.size, ab[k])) for k in range(3)]
        self.a, self.b = a.ravel(), b.ravel()

    def hermite_block(self, ca, cb):
        """Hermite coefficients E_tuv (tmax+1, umax+1, vmax+1, nprim) of the component pair."""
        (ia, ja, ka), (ib, jb, kb) = ca, cb
        ex = self.e[0][ia, ib, :ia + ib + 1]
        ey = self.e[1][ja, jb, :ja + jb + 1]
        ez = self.e[2][ka, kb, :ka + kb + 1]
        return ex[:, None, None, :] * ey[None, :, None, :] * ez[None, None, :, :]


def _angular_factor(c):
    """Axis-aligned norm ratio of component (i, j, k): sqrt((2l-1)!! / ((2i-1)!! (2j-1)!! (2k-1)!!))."""
    i, j, k = c
    l = i + j + k
    return math.sqrt(_dfact(2 * l - 1) / (_dfact(2 * i - 1) * _dfact(2 * j - 1) * _dfact(2 * k - 1)))


def _aux_pair(sc):
    """Pair data of an auxiliary shell with a zero-exponent s partner (the three-/two-centre limit)."""
    pr = _Pair.__new__(_Pair)
    pr.la, pr.lb = sc.l, 0
    pr.p = sc.exps.copy()
    pr.pc = np.repeat(sc.centre[None, :], len(sc.exps), axis=0)
    pr.w = sc.coefs * _prim_norm(sc.exps, sc.l)
    pr.e = [hermite_e(sc.l, 0, sc.exps, np.zeros_like(sc.exps), np.zeros_like(sc.exps)) for _ in range(3)]
    pr.a, pr.b = sc.exps, np.zeros_like(sc.exps)
    return pr


def _offsets(shells):
    off, n = [], 0
    for s in shells:
        off.append(n)
        n += s.size
    return off, n


# ---- one-electron integrals ------------------------------------------------------------------------------
def one_electron(shells, charges, positions):
    """(S, T, V) over the Cartesian functions of `shells` (unnormalized contractions; see normalize)."""
    off, n = _offsets(shells)
    s_m, t_m, v_m = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    charges = np.asarray(charges, float)
    positions = np.asarray(positions, float).reshape(-1, 3)
    for ia, sa in enumerate(shells):
        for ib in range(ia, len(shells)):
            sb = shells[ib]
            pr = _Pair(sa, sb)
            # the kinetic integral needs the overlap factors with the b power raised by 2
            e2 = [hermite_e(sa.l, sb.l + 2, pr.a, pr.b, np.full(pr.p.size, (sa.centre - sb.centre)[k]))
                  for k in range(3)]
            pref = (np.pi / pr.p) ** 1.5
            rs = [hermite_r(sa.l + sb.l, pr.p, pr.pc - c[None, :]) for c in positions]
            for x, ca in enumerate(cart_components(sa.l)):
                for y, cb in enumerate(cart_components(sb.l)):
                    fa, fb = _angular_factor(ca), _angular_factor(cb)
                    sx = [pr.e[k][ca[k], cb[k], 0] for k in range(3)]
                    sval = np.sum(pr.w * pref * sx[0] * sx[1] * sx[2])

                    def s1(k, j):  # overlap factor along axis k with the b exponent shifted to j
                        if j < 0:
                            return np.zeros_like(pr.p)
                        return e2[k][ca[k], j, 0]

                    tk = 0.0
                    for k in range(3):
                        j = cb[k]
                        d2 = j * (j - 1) * s1(k, j - 2) - 2.0 * pr.b * (2 * j + 1) * s1(k, j) + 4.0 * pr.b ** 2 * s1(k, j + 2)
                        others = np.prod([sx[m] for m in range(3) if m != k], axis=0)
                        tk = tk + d2 * others
                    tval = -0.5 * np.sum(pr.w * pref * tk)
                    eb = pr.hermite_block(ca, cb)
                    vval = 0.0
                    for zc, r in zip(charges, rs):
                        tt, uu, vv = eb.shape[:3]
                        vval -= zc * np.sum(pr.w * 2.0 * np.pi / pr.p * np.sum(eb * r[:tt, :uu, :vv], axis=(0, 1, 2)))
                    i, j = off[ia] + x, off[ib] + y
                    s_m[i, j] = s_m[j, i] = sval * fa * fb
                    t_m[i, j] = t_m[j, i] = tval * fa * fb
                    v_m[i, j] = v_m[j, i] = vval * fa * fb
    return s_m, t_m, v_m


# ---- Coulomb integrals ---------------------------------------------------------------------------------
def _coulomb_pairs(pab, pcd, cab, ccd):
    """(ab|cd) for two primitive-pair sets and one component pair each."""
    e1 = pab.hermite_block(*cab)  # (t, u, v, n1)
    e2 = pcd.hermite_block(*ccd)  # (tau, nu, phi, n2)
    t1, u1, v1 = e1.shape[:3]
    t2, u2, v2 = e2.shape[:3]
    lmax = (t1 - 1) + (u1 - 1) + (v1 - 1) + (t2 - 1) + (u2 - 1) + (v2 - 1)
    p, q = pab.p[:, None], pcd.p[None, :]
    alpha = p * q / (p + q)
    pq = pab.pc[:, None, :] - pcd.pc[None, :, :]
    r = hermite_r(lmax, alpha, pq)  # (L+1, L+1, L+1, n1, n2)
    sign2 = np.array([(-1.0) ** k for k in range(max(t2, u2, v2))])
    e2s = e2 * sign2[:t2, None, None, None] * sign2[None, :u2, None, None] * sign2[None, None, :v2, None]
    tot = np.zeros((e1.shape[3], e2.shape[3]))
    for t in range(t1):
        for u in range(u1):
            for v in range(v1):
                blk = r[t:t + t2, u:u + u2, v:v + v2]  # (t2, u2, v2, n1, n2)
                tot += e1[t, u, v][:, None] * np.einsum("abcj,abcij->ij", e2s, blk)
    pref = 2.0 * np.pi ** 2.5 / (p * q * np.sqrt(p + q))
    return float(np.sum(pab.w[:, None] * pcd.w[None, :] * pref * tot))


def three_center(shells, aux_shells):
    """raw[P, mu, nu] = (mu nu | P) over Cartesian functions (unnormalized contractions)."""
    off, n = _offsets(shells)
    aoff, naux = _offsets(aux_shells)
    out = np.zeros((naux, n, n))
    aux_pairs = [_aux_pair(s) for s in aux_shells]
    for ia, sa in enumerate(shells):
        for ib in range(ia, len(shells)):
            pr = _Pair(sa, shells[ib])
            for x, ca in enumerate(cart_components(sa.l)):
                for y, cb in enumerate(cart_components(shells[ib].l)):
                    f_ab = _angular_factor(ca) * _angular_factor(cb)
                    for ic, sc in enumerate(aux_shells):
                        for z, cc in enumerate(cart_components(sc.l)):
                            v = _coulomb_pairs(pr, aux_pairs[ic], (ca, cb), (cc, (0, 0, 0))) * f_ab * _angular_factor(cc)
                            out[aoff[ic] + z, off[ia] + x, off[ib] + y] = v
                            out[aoff[ic] + z, off[ib] + y, off[ia] + x] = v
    return out


def three_center_prim(shells, aux_shells):
    """raw[P, mu, nu] = (mu nu | P) as three_center, for shells and auxiliary shells of one primitive
    each, vectorized over the auxiliary shells of each angular momentum: per orbital shell pair and
    auxiliary l, one Hermite-integral table R_tuv over all those auxiliary centres."""
    if any(len(s.exps) != 1 for s in shells) or any(len(s.exps) != 1 for s in aux_shells):
        raise ValueError("three_center_prim: every shell must have exactly one primitive")
    off, n = _offsets(shells)
    aoff, naux = _offsets(aux_shells)
    out = np.zeros((naux, n, n))
    groups = {}
    for ic, sc in enumerate(aux_shells):
        groups.setdefault(sc.l, []).append(ic)
    gdata = []
    for lc, idx in sorted(groups.items()):
        g = np.array([aux_shells[i].exps[0] for i in idx])
        cc = np.array([aux_shells[i].centre for i in idx])
        w = np.array([aux_shells[i].coefs[0] for i in idx]) * _prim_norm(g, lc)
        ec = hermite_e(lc, 0, g, np.zeros_like(g), np.zeros_like(g))  # [i, 0, t, ng]
        comps = cart_components(lc)
        # E^c with the (-1)^(tau+nu+phi) sign of the ket and the component's angular factor
        e_c = np.zeros((len(comps), lc + 1, lc + 1, lc + 1, len(g)))
        for z, (i, j, k) in enumerate(comps):
            blk = ec[i, 0, :i + 1][:, None, None, :] * ec[j, 0, :j + 1][None, :, None, :] * ec[k, 0, :k + 1][None, None, :, :]
            sg = (-1.0) ** (np.arange(i + 1)[:, None, None] + np.arange(j + 1)[None, :, None] + np.arange(k + 1)[None, None, :])
            e_c[z, :i + 1, :j + 1, :k + 1] = blk * sg[..., None] * _angular_factor((i, j, k))
        rows = np.array([aoff[i] for i in idx])[:, None] + np.arange(len(comps))[None, :]  # (ng, nc)
        gdata.append((lc, g, cc, w, e_c, rows))
    for ia, sa in enumerate(shells):
        ca_list = cart_components(sa.l)
        for ib in range(ia, len(shells)):
            sb = shells[ib]
            cb_list = cart_components(sb.l)
            a, b = sa.exps[0], sb.exps[0]
            p = a + b
            pc = (a * sa.centre + b * sb.centre) / p
            wab = sa.coefs[0] * _prim_norm(a, sa.l) * sb.coefs[0] * _prim_norm(b, sb.l)
            ex = [hermite_e(sa.l, sb.l, a, b, sa.centre[k] - sb.centre[k]) for k in range(3)]
            t1 = sa.l + sb.l + 1
            e_ab = np.zeros((len(ca_list), len(cb_list), t1, t1, t1))
            for x, ca in enumerate(ca_list):
                for y, cb in enumerate(cb_list):
                    e_ab[x, y, :ca[0] + cb[0] + 1, :ca[1] + cb[1] + 1, :ca[2] + cb[2] + 1] = (
                        ex[0][ca[0], cb[0], :ca[0] + cb[0] + 1][:, None, None]
                        * ex[1][ca[1], cb[1], :ca[1] + cb[1] + 1][None, :, None]
                        * ex[2][ca[2], cb[2], :ca[2] + cb[2] + 1][None, None, :]
                        * _angular_factor(ca) * _angular_factor(cb))
            for lc, g, cc, w, e_c, rows in gdata:
                lmax = t1 - 1 + lc
                alpha = p * g / (p + g)
                r = hermite_r(lmax, alpha, pc[None, :] - cc)  # (L+1, L+1, L+1, ng)
                acc = np.zeros((t1, t1, t1, e_c.shape[0], len(g)))
                for tau in range(lc + 1):
                    for nu in range(lc + 1 - tau):
                        for phi in range(lc + 1 - tau - nu):
                            acc += e_c[:, tau, nu, phi, :][None, None, None] * r[tau:tau + t1, nu:nu + t1, phi:phi + t1, None, :]
                pref = 2.0 * np.pi ** 2.5 / (p * g * np.sqrt(p + g)) * wab * w
                val = np.einsum("xytuv,tuvcg->gcxy", e_ab, acc, optimize=True) * pref[:, None, None, None]
                cols_a = off[ia] + np.arange(len(ca_list))
                cols_b = off[ib] + np.arange(len(cb_list))
                out[rows[:, :, None, None], cols_a[None, None, :, None], cols_b[None, None, None, :]] = val
                out[rows[:, :, None, None], cols_b[None, None, None, :], cols_a[None, None, :, None]] = val
    return out


def _prim_group_data(aux_shells, sign):
    """Per angular momentum, the single-primitive shells' exponents, centres, weights, Hermite
    coefficients E_tuv of each component about its own centre (times (-1)^(t+u+v) when `sign`, the
    ket side of a Coulomb integral, and the component's angular factor) and output rows."""
    aoff, _ = _offsets(aux_shells)
    groups = {}
    for ic, sc in enumerate(aux_shells):
        groups.setdefault(sc.l, []).append(ic)
    out = []
    for lc, idx in sorted(groups.items()):
        g = np.array([aux_shells[i].exps[0] for i in idx])
        cc = np.array([aux_shells[i].centre for i in idx])
        w = np.array([aux_shells[i].coefs[0] for i in idx]) * _prim_norm(g, lc)
        ec = hermite_e(lc, 0, g, np.zeros_like(g), np.zeros_like(g))
        comps = cart_components(lc)
        e_c = np.zeros((len(comps), lc + 1, lc + 1, lc + 1, len(g)))
        for z, (i, j, k) in enumerate(comps):
            blk = ec[i, 0, :i + 1][:, None, None, :] * ec[j, 0, :j + 1][None, :, None, :] * ec[k, 0, :k + 1][None, None, :, :]
            if sign:
                blk = blk * ((-1.0) ** (np.arange(i + 1)[:, None, None] + np.arange(j + 1)[None, :, None]
                                        + np.arange(k + 1)[None, None, :]))[..., None]
            e_c[z, :i + 1, :j + 1, :k + 1] = blk * _angular_factor((i, j, k))
        rows = np.array([aoff[i] for i in idx])[:, None] + np.arange(len(comps))[None, :]
        out.append((lc, idx, g, cc, w, e_c, rows))
    return out


def two_center_prim(aux_shells):
    """(P|Q) as two_center, for single-primitive shells, vectorized over the Q shells of each angular
    momentum."""
    if any(len(s.exps) != 1 for s in aux_shells):
        raise ValueError("two_center_prim: every shell must have exactly one primitive")
    aoff, naux = _offsets(aux_shells)
    out = np.zeros((naux, naux))
    bra = _prim_group_data(aux_shells, False)
    ket = _prim_group_data(aux_shells, True)
    for la, idx_a, ga, ca, wa, ea, rows_a in bra:
        t1 = la + 1
        for ia in range(len(idx_a)):
            a, ctr = ga[ia], ca[ia]
            for lb, idx_b, gb, cb, wb, eb, rows_b in ket:
                lmax = la + lb
                alpha = a * gb / (a + gb)
                r = hermite_r(lmax, alpha, ctr[None, :] - cb)
                acc = np.zeros((t1, t1, t1, eb.shape[0], len(gb)))
                for tau in range(lb + 1):
                    for nu in range(lb + 1 - tau):
                        for phi in range(lb + 1 - tau - nu):
                            acc += eb[:, tau, nu, phi, :][None, None, None] * r[tau:tau + t1, nu:nu + t1, phi:phi + t1, None, :]
                pref = 2.0 * np.pi ** 2.5 / (a * gb * np.sqrt(a + gb)) * wa[ia] * wb
                val = np.einsum("xtuv,tuvcg->xgc", ea[:, :, :, :, ia], acc, optimize=True) * pref[None, :, None]
                out[rows_a[ia][:, None, None], rows_b[None, :, :]] = val
    return out


def self_overlap_prim(shells):
    """Diagonal of the overlap of single-primitive shells: coef^2 for every component (a unit
    coefficient primitive shell's components are normalized, the angular factor makes each component
    x^i y^j z^k carry the norm of x^l)."""
    if any(len(s.exps) != 1 for s in shells):
        raise ValueError("self_overlap_prim: every shell must have exactly one primitive")
    return np.concatenate([np.full(s.size, float(s.coefs[0]) ** 2) for s in shells])


def nabla(shells):
    """D[k][mu, nu] = <mu | d/dx_k | nu> over the Cartesian functions of `shells` (unnormalized
    contractions): d/dx x^j exp(-b x^2) = j x^(j-1) exp(-b x^2) - 2 b x^(j+1) exp(-b x^2) turns each
    axis factor into overlap factors with the ket power lowered and raised by one."""
    off, n = _offsets(shells)
    d = np.zeros((3, n, n))
    for ia, sa in enumerate(shells):
        for ib, sb in enumerate(shells):
            pr = _Pair(sa, sb)
            e1 = [hermite_e(sa.l, sb.l + 1, pr.a, pr.b, np.full(pr.p.size, (sa.centre - sb.centre)[k])) for k in range(3)]
            pref = pr.w * (np.pi / pr.p) ** 1.5
            for x, ca in enumerate(cart_components(sa.l)):
                for y, cb in enumerate(cart_components(sb.l)):
                    f = _angular_factor(ca) * _angular_factor(cb)
                    s0 = [e1[k][ca[k], cb[k], 0] for k in range(3)]
                    for k in range(3):
                        j = cb[k]
                        dk = -2.0 * pr.b * e1[k][ca[k], j + 1, 0]
                        if j > 0:
                            dk = dk + j * e1[k][ca[k], j - 1, 0]
                        others = np.prod([s0[m] for m in range(3) if m != k], axis=0)
                        d[k, off[ia] + x, off[ib] + y] = f * float(np.sum(pref * dk * others))
    return d


def two_center(aux_shells):
    """(P|Q) over the Cartesian auxiliary functions (unnormalized contractions)."""
    aoff, naux = _offsets(aux_shells)
    out = np.zeros((naux, naux))
    pairs = [_aux_pair(s) for s in aux_shells]
    for ia, sa in enumerate(aux_shells):
        for ib in range(ia, len(aux_shells)):
            sb = aux_shells[ib]
            for x, ca in enumerate(cart_components(sa.l)):
                for y, cb in enumerate(cart_components(sb.l)):
                    v = _coulomb_pairs(pairs[ia], pairs[ib], (ca, (0, 0, 0)), (cb, (0, 0, 0)))
                    v *= _angular_factor(ca) * _angular_factor(cb)
                    out[aoff[ia] + x, aoff[ib] + y] = out[aoff[ib] + y, aoff[ia] + x] = v
    return out


def four_center(shells):
    """(mu nu | lam sig) over Cartesian functions, chemist notation (small bases only: O(n^4)); each
    unique element of the eight-fold permutational symmetry is computed once."""
    off, n = _offsets(shells)
    out = np.zeros((n, n, n, n))
    comps = [(si, c) for si, s in enumerate(shells) for c in cart_components(s.l)]
    pairs = {}
    for ia in range(len(shells)):
        for ib in range(ia + 1):
            pairs[ia, ib] = _Pair(shells[ia], shells[ib])
    fac = [_angular_factor(c) for _, c in comps]
    for a in range(n):
        for b in range(a + 1):
            ab = a * (a + 1) // 2 + b
            for c in range(n):
                for d in range(c + 1):
                    if c * (c + 1) // 2 + d > ab:
                        continue
                    (sa, ca), (sb, cb), (sc, cc), (sd, cd) = comps[a], comps[b], comps[c], comps[d]
                    v = _coulomb_pairs(pairs[sa, sb], pairs[sc, sd], (ca, cb), (cc, cd)) * fac[a] * fac[b] * fac[c] * fac[d]
                    for (i, j) in ((a, b), (b, a)):
                        for (k, m) in ((c, d), (d, c)):
                            out[i, j, k, m] = out[k, m, i, j] = v
    return out


def normalization(s_diag):
    """Per-function scale that makes every Cartesian function unit-normalized (normalizeCartesianBasis)."""
    return 1.0 / np.sqrt(s_diag)


def nuclear_repulsion(charges, positions):
    z = np.asarray(charges, float)
    r = np.asarray(positions, float).reshape(-1, 3)
    e = 0.0
    for i in range(len(z)):
        for j in range(i):
            e += z[i] * z[j] / np.linalg.norm(r[i] - r[j])
    return e


# ---- automatic auxiliary basis (AuxiliaryBasis.cpp:40-190) ------------------------------------------------
def _lval(z):
    return 0 if z <= 2 else (1 if z <= 18 else (2 if z <= 54 else 3))


def auto_aux_shells(shells, kind="PAUTO", f_sam=1.5, lmax_inc=1):
    """The PAUTO / AUTO auxiliary shells of an orbital basis (uncontracted, coefficient 1). Per centre
    (in order of centre coordinates): the candidate (exponent, l) pairs -- every ordered pair of the
    centre's unique shells and of their primitives, exponents and l summed (PAUTO), or each primitive
    doubled (AUTO) -- are clustered greedily from the largest exponent: a cluster takes every remaining
    candidate within the ratio f_sam, its exponent is the geometric mean of its unique (exponent, l)
    pairs, its l is min(max(l in cluster, running l), lmax_abs) with lmax_abs = max(lmax_obs +
    lmax_inc, 2 lval(Z)), and it emits one shell per l = 0..that l."""
    centres = {}
    for s in shells:
        key = tuple(float(v) for v in s.centre)
        centres.setdefault(key, {"element": s.element, "shells": {}})
        centres[key]["shells"][(s.l, tuple(float(e) for e in s.exps))] = s
    out = []
    for key in sorted(centres):
        info = centres[key]
        uniq = sorted(info["shells"].keys())
        lmax_obs = max(l for l, _ in uniq)
        lmax_abs = max(lmax_obs + lmax_inc, 2 * _lval(ELEMENTS[info["element"]]))
        cand = []
        if kind == "PAUTO":
            for li, ei in uniq:
                for a in ei:
                    for lj, ej in uniq:
                        for b in ej:
                            cand.append((a + b, li + lj))
        else:
            for li, ei in uniq:
                for a in ei:
                    cand.append((2.0 * a, 2 * li))
        remaining = [True] * len(cand)
        left = len(cand)
        am_current = 0
        while left > 0:
            idx, best = 0, -1.0
            for k, (e, _) in enumerate(cand):
                if remaining[k] and e > best:
                    best, idx = e, k
            ecur = cand[idx][0]
            trial = [idx]
            remaining[idx] = False
            left -= 1
            for k, (e, _) in enumerate(cand):
                if remaining[k] and ecur / e < f_sam:
                    trial.append(k)
                    remaining[k] = False
                    left -= 1
            unique = []
            for t in trial:
                e, l = cand[t]
                if not any(u[1] == l and abs(u[0] - e) < 1e-6 for u in unique):
                    unique.append((e, l))
            esel = math.exp(sum(math.log(u[0]) for u in unique) / len(unique))
            am_sel = min(max(max(cand[t][1] for t in trial), am_current), lmax_abs)
            am_current = am_sel
            for l in range(am_sel + 1):
                out.append(Shell(np.array(key), l, [esel], [1.0], info["element"]))
    return out


class GaussianMolecule:
    """A molecule in a Gaussian basis (Cartesian functions, each normalized) with its automatic RI
    auxiliary basis: the FP64 inputs of the NON_REL RI SCF, prepared as the reference prepares them
    (src/main.cpp:1068-1090): PAUTO (or AUTO) auxiliary shells, per-function normalization, the raw
    three-centre tensor (mu nu|P), the metric (P|Q) orthogonalized canonically with threshold 1e-10
    (directions below it dropped), B = X^T raw.

    atoms: [(symbol, xyz in bohr)]; basis: parse_gbs result."""

    def __init__(self, atoms, basis, n_electrons=None, aux_kind="PAUTO", metric_threshold=1e-10):
        self.symbols = [a[0] for a in atoms]
        self.charges = np.array([float(ELEMENTS[a[0]]) for a in atoms])
        self.positions = np.array([a[1] for a in atoms], dtype=np.float64).reshape(-1, 3)
        self.shells = molecule_shells(list(zip(self.symbols, self.positions)), basis)
        self.aux_shells = auto_aux_shells(self.shells, aux_kind)
        self.metric_threshold = metric_threshold
        self.n_electrons = int(round(self.charges.sum())) if n_electrons is None else n_electrons
        self._ints = None

    @property
    def n(self):
        return sum(s.size for s in self.shells)

    @property
    def n_aux(self):
        return self.integrals()["eri3"].shape[0]

    def integrals(self):
        """dict: S, T, V, h, eri3 ((n_aux_kept, n, n)), e_nuc, n_aux_raw, n_aux_dropped."""
        if self._ints is None:
            s, t, v = one_electron(self.shells, self.charges, self.positions)
            nrm = normalization(np.diag(s))
            o2 = np.outer(nrm, nrm)
            sa, _, _ = one_electron(self.aux_shells, [], np.zeros((0, 3)))
            an = normalization(np.diag(sa))
            raw = three_center(self.shells, self.aux_shells) * an[:, None, None] * o2[None]
            metric = two_center(self.aux_shells) * np.outer(an, an)
            w, u = np.linalg.eigh(metric)
            keep = w > self.metric_threshold
            x = u[:, keep] / np.sqrt(w[keep])
            b = np.einsum("qmn,qp->pmn", raw, x, optimize=True)
            b = 0.5 * (b + np.transpose(b, (0, 2, 1)))
            self._ints = {"S": s * o2, "T": t * o2, "V": v * o2, "h": (t + v) * o2, "eri3": np.ascontiguousarray(b),
                          "e_nuc": nuclear_repulsion(self.charges, self.positions), "n_aux_raw": metric.shape[0],
                          "n_aux_dropped": int(np.sum(~keep))}
        return dict(self._ints)
