"""Small molecules in s-type Gaussian bases: the AO and RI integrals that feed the NON_REL SCF.

The reference takes its integrals from libcint (AO_ints/ThreeCenterIntegrals.cpp:18-21), which is
not available here. This module is a closed-form engine for s-type Gaussians only -- a model
chemistry that is enough to produce families of real SCF problems (overlap, kinetic, nuclear
attraction, RI three-centre (mu nu|P) and two-centre (P|Q) Coulomb integrals) on seeded, perturbed
geometries. It is host FP64 code, like the integral generation of the reference.

Primitive integrals of g_A(r) = exp(-a |r - A|^2) (Szabo and Ostlund, "Modern Quantum Chemistry",
appendix A), with p = a + b, mu = a b / p, P = (a A + b B) / p:
  S    = (pi/p)^(3/2) exp(-mu |A-B|^2)
  T    = mu (3 - 2 mu |A-B|^2) S
  V_C  = -Z_C 2 pi / p exp(-mu |A-B|^2) F0(p |P-C|^2)
  (ab|cd) = 2 pi^(5/2) / (p q sqrt(p+q)) exp(-mu_ab |A-B|^2 - mu_cd |C-D|^2) F0(p q/(p+q) |P-Q|^2)
with F0(t) = erf(sqrt t) sqrt(pi/t) / 2. The three-centre (ab|c) and two-centre (c|d) integrals
are the limits of (ab|cd) in which the missing functions have exponent 0 (a constant factor 1).
Contracted basis functions are normalized to <g|g> = 1; auxiliary functions are primitives with the
same (2a/pi)^(3/4) normalization (the fitted tensor does not depend on it).

The RI tensor handed to the Fock build is B_P = sum_Q (mu nu|Q) [(P|Q)^(-1/2)]_QP, the layout of
AO_ints/ThreeCenterIntegrals.h:42-70 ((n_aux, n*n), row P the n x n block).
"""
import numpy as np
from scipy.special import erf

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903

# STO-3G hydrogen 1s for zeta = 1 (Hehre, Stewart and Pople 1969); exponents scale with zeta^2.
STO3G_1S_EXP = np.array([0.109818, 0.405771, 2.22766])
STO3G_1S_COEF = np.array([0.444635, 0.535328, 0.154329])


def boys0(t):
    t = np.asarray(t, dtype=np.float64)
    out = np.ones_like(t)
    big = t > 1e-12
    st = np.sqrt(t[big])
    out[big] = 0.5 * np.sqrt(np.pi) * erf(st) / st
    small = ~big
    out[small] = 1.0 - t[small] / 3.0
    return out


def _norm(a):
    return (2.0 * a / np.pi) ** 0.75


class Basis:
    """Contracted s functions: each is (centre xyz, exponents, coefficients of normalized primitives)."""

    def __init__(self, functions):
        self.centres = np.array([f[0] for f in functions], dtype=np.float64).reshape(-1, 3)
        self.exps = [np.asarray(f[1], dtype=np.float64) for f in functions]
        coefs = [np.asarray(f[2], dtype=np.float64) * _norm(np.asarray(f[1], dtype=np.float64)) for f in functions]
        # normalize each contraction
        self.coefs = []
        for c, e, r in zip(coefs, self.exps, self.centres):
            s = _prim_overlap(e[:, None], r, e[None, :], r)
            self.coefs.append(c / np.sqrt(c @ s @ c))

    def __len__(self):
        return len(self.exps)

    def primitives(self):
        """Flattened primitive arrays (function index, exponent, coefficient, centre)."""
        fi, ex, co, ce = [], [], [], []
        for i, (e, c) in enumerate(zip(self.exps, self.coefs)):
            fi += [i] * len(e)
            ex += list(e)
            co += list(c)
            ce += [self.centres[i]] * len(e)
        return np.array(fi), np.array(ex), np.array(co), np.array(ce).reshape(-1, 3)


def _prim_overlap(a, ra, b, rb):
    p = a + b
    d2 = np.sum((np.asarray(ra) - np.asarray(rb)) ** 2, axis=-1)
    return (np.pi / p) ** 1.5 * np.exp(-a * b / p * d2)


def _contract(fi, val, n, fj=None):
    """Sum primitive-pair values into contracted functions."""
    fj = fi if fj is None else fj
    out = np.zeros((n, n))
    np.add.at(out, (fi[:, None], fj[None, :]), val)
    return out


def one_electron(basis, charges, positions):
    """Overlap S, kinetic T and nuclear attraction V (n x n) of an s basis."""
    fi, ex, co, ce = basis.primitives()
    a, b = ex[:, None], ex[None, :]
    p = a + b
    mu = a * b / p
    d2 = np.sum((ce[:, None, :] - ce[None, :, :]) ** 2, axis=-1)
    cc = co[:, None] * co[None, :]
    s = (np.pi / p) ** 1.5 * np.exp(-mu * d2)
    t = mu * (3.0 - 2.0 * mu * d2) * s
    pc = (a[..., None] * ce[:, None, :] + b[..., None] * ce[None, :, :]) / p[..., None]
    v = np.zeros_like(s)
    for z, r in zip(charges, positions):
        v -= z * 2.0 * np.pi / p * np.exp(-mu * d2) * boys0(p * np.sum((pc - r) ** 2, axis=-1))
    n = len(basis)
    return _contract(fi, cc * s, n), _contract(fi, cc * t, n), _contract(fi, cc * v, n)


def _pair_data(basis):
    fi, ex, co, ce = basis.primitives()
    a, b = ex[:, None], ex[None, :]
    p = a + b
    mu = a * b / p
    d2 = np.sum((ce[:, None, :] - ce[None, :, :]) ** 2, axis=-1)
    kab = co[:, None] * co[None, :] * np.exp(-mu * d2)
    pc = (a[..., None] * ce[:, None, :] + b[..., None] * ce[None, :, :]) / p[..., None]
    return fi, p, kab, pc


def four_center(basis):
    """Exact (ab|cd) over contracted functions, chemists' notation, shape (n, n, n, n). Small bases only."""
    fi, p, kab, pc = _pair_data(basis)
    n = len(basis)
    npr = len(fi)
    pf, kf, cf = p.reshape(-1), kab.reshape(-1), pc.reshape(-1, 3)
    q = pf[None, :]
    pp = pf[:, None]
    d2 = np.sum((cf[:, None, :] - cf[None, :, :]) ** 2, axis=-1)
    val = 2.0 * np.pi ** 2.5 / (pp * q * np.sqrt(pp + q)) * kf[:, None] * kf[None, :] * boys0(pp * q / (pp + q) * d2)
    val = val.reshape(npr, npr, npr, npr)
    out = np.zeros((n, n, n, n))
    np.add.at(out, (fi[:, None, None, None], fi[None, :, None, None], fi[None, None, :, None], fi[None, None, None, :]),
              val)
    return out


def three_center(basis, aux_exps, aux_centres):
    """(mu nu|P) for primitive normalized s auxiliary functions, shape (n_aux, n, n)."""
    fi, p, kab, pc = _pair_data(basis)
    g = np.asarray(aux_exps, dtype=np.float64)
    gc = np.asarray(aux_centres, dtype=np.float64).reshape(-1, 3)
    n = len(basis)
    out = np.empty((len(g), n, n))
    for k, (gam, c) in enumerate(zip(g, gc)):
        d2 = np.sum((pc - c) ** 2, axis=-1)
        val = 2.0 * np.pi ** 2.5 / (p * gam * np.sqrt(p + gam)) * kab * boys0(p * gam / (p + gam) * d2) * _norm(gam)
        out[k] = _contract(fi, val, n)
    return out


def two_center(aux_exps, aux_centres):
    """(P|Q) Coulomb metric of primitive normalized s auxiliary functions."""
    g = np.asarray(aux_exps, dtype=np.float64)
    gc = np.asarray(aux_centres, dtype=np.float64).reshape(-1, 3)
    a, b = g[:, None], g[None, :]
    d2 = np.sum((gc[:, None, :] - gc[None, :, :]) ** 2, axis=-1)
    return 2.0 * np.pi ** 2.5 / (a * b * np.sqrt(a + b)) * boys0(a * b / (a + b) * d2) * _norm(a) * _norm(b)


def inverse_sqrt(m, threshold=1e-10):
    """Symmetric M^(-1/2), eigenvalues below threshold * max dropped."""
    w, v = np.linalg.eigh(m)
    keep = w > threshold * w[-1]
    return (v[:, keep] / np.sqrt(w[keep])) @ v[:, keep].T


def nuclear_repulsion(charges, positions):
    e = 0.0
    for i in range(len(charges)):
        for j in range(i):
            e += charges[i] * charges[j] / np.linalg.norm(positions[i] - positions[j])
    return e


def even_tempered(alpha0, beta, count):
    return alpha0 * beta ** np.arange(count)


class Molecule:
    """Nuclei (charges, positions in bohr), an orbital basis and an s auxiliary basis.

    `basis_spec` / `aux_spec` map a nuclear charge to a list of (exponents, coefficients) contractions
    (orbital basis) or to an array of auxiliary exponents."""

    def __init__(self, charges, positions, basis_spec, aux_spec, n_electrons=None):
        self.charges = np.asarray(charges, dtype=np.float64)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        funcs = []
        for z, r in zip(self.charges, self.positions):
            for e, c in basis_spec[int(z)]:
                funcs.append((r, e, c))
        self.basis = Basis(funcs)
        ae, ac = [], []
        for z, r in zip(self.charges, self.positions):
            for e in aux_spec[int(z)]:
                ae.append(e)
                ac.append(r)
        self.aux_exps = np.array(ae)
        self.aux_centres = np.array(ac).reshape(-1, 3)
        self.n_electrons = int(round(self.charges.sum())) if n_electrons is None else n_electrons

    @property
    def n(self):
        return len(self.basis)

    @property
    def n_aux(self):
        return len(self.aux_exps)

    def integrals(self):
        """dict: S, T, V, h (= T + V), eri3 (B tensor, (n_aux, n, n)), e_nuc."""
        s, t, v = one_electron(self.basis, self.charges, self.positions)
        raw = three_center(self.basis, self.aux_exps, self.aux_centres)
        metric = two_center(self.aux_exps, self.aux_centres)
        x = inverse_sqrt(metric)
        b = np.einsum("qmn,qp->pmn", raw, x, optimize=True)
        b = 0.5 * (b + np.transpose(b, (0, 2, 1)))
        return {"S": s, "T": t, "V": v, "h": t + v, "eri3": np.ascontiguousarray(b),
                "e_nuc": nuclear_repulsion(self.charges, self.positions)}


# Water in an s-only even-tempered basis (model chemistry): 10 s functions on O, 5 on each H;
# auxiliary s functions 18 on O, 10 on each H.
WATER_CHARGES = np.array([8.0, 1.0, 1.0])
WATER_GEOMETRY = np.array([[0.0, 0.0, 0.0], [1.4305, 1.1078, 0.0], [-1.4305, 1.1078, 0.0]])
WATER_BASIS = {8: [([e], [1.0]) for e in even_tempered(0.15, 2.8, 10)],
               1: [([e], [1.0]) for e in even_tempered(0.08, 3.0, 5)]}
WATER_AUX = {8: even_tempered(0.3, 2.2, 18), 1: even_tempered(0.16, 2.5, 10)}


def perturbed_geometries(base, count, amplitude, seed):
    """`count` geometries: base + amplitude * N(0, 1) per coordinate (bohr), seeded."""
    g = np.random.Generator(np.random.PCG64(seed))
    return [base + amplitude * g.standard_normal(base.shape) for _ in range(count)]


def water(positions=None):
    return Molecule(WATER_CHARGES, WATER_GEOMETRY if positions is None else positions, WATER_BASIS, WATER_AUX)
