"""The Gaussian integral engine and the automatic auxiliary basis (rerdmft_amd/integrals.py), on the CPU.

Pins, each independent of the McMurchie-Davidson recursions under test:
- s functions: the closed-form s-type integrals of rerdmft_amd/molecule.py (themselves checked against
  the published H2/STO-3G values);
- p, d and f functions: derivatives of lower-l integrals with respect to a function's centre (central
  differences): d/dA_x exp(-a|r-A|^2) = 2a (x - A_x) exp(-a|r-A|^2), so (p_x ...) = (1/2a) d/dA_x (s ...),
  (d_xy ...) = (1/4a^2) d2/dA_x dA_y (s ...), (d_xx ...) = (1/4a^2) (d2/dA_x^2 + 2a) (s ...), for overlap,
  kinetic, nuclear attraction, three-centre and two-centre Coulomb integrals;
- end to end, the reference's own energies (examples/reference.yaml): RI NON_REL HF of LiH/6-31G with
  the PAUTO auxiliary basis, -7.9793032232 Ha (:990; basis tests/golden/lih-6-31g.gbs, the reference's
  examples/lih-6-31g.gbs), through the oracle's nonRelFockMatrix RI build; and the dense HF energies of
  Ne/cc-pVDZ, Cartesian -128.488866172 Ha (:2001) and spherical -128.4887755517 Ha (:2027), which
  exercise d shells in all four-centre integrals. Tolerance 1e-8 Ha.
"""
import os

import numpy as np
import pytest

from oracle import oracle
from rerdmft_amd import integrals as I
from rerdmft_amd import molecule, scf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def prim(centre, l, a):
    """Single normalized primitive shell."""
    return I.Shell(np.asarray(centre, float), l, [a], [1.0])


def raw_scale(l, a, comp):
    """Factor that turns the integral of a normalized component into that of x^i y^j z^k exp(-a r^2)."""
    return 1.0 / (I._prim_norm(a, l) * I._angular_factor(comp))


def test_s_functions_match_the_closed_forms():
    mol = molecule.water()
    shells = [I.Shell(r, 0, e, c) for z, r in zip(mol.charges, mol.positions) for e, c in molecule.WATER_BASIS[int(z)]]
    aux = [I.Shell(r, 0, [e], [1.0]) for e, r in zip(mol.aux_exps, mol.aux_centres)]
    s, t, v = I.one_electron(shells, mol.charges, mol.positions)
    nrm = I.normalization(np.diag(s))
    s0, t0, v0 = molecule.one_electron(mol.basis, mol.charges, mol.positions)
    o2 = np.outer(nrm, nrm)
    assert np.max(np.abs(s * o2 - s0)) < 1e-13 and np.max(np.abs(t * o2 - t0)) < 1e-11
    assert np.max(np.abs(v * o2 - v0)) < 1e-11
    raw = I.three_center(shells, aux) * o2[None]
    assert np.max(np.abs(raw - molecule.three_center(mol.basis, mol.aux_exps, mol.aux_centres))) < 1e-12
    assert np.max(np.abs(I.two_center(aux) - molecule.two_center(mol.aux_exps, mol.aux_centres))) < 1e-11


def _fd(f, centre, axes, h=1e-4):
    """Mixed central difference of f(centre) along one or two axes."""
    def at(dx):
        c = np.array(centre, float)
        for ax, d in dx:
            c[ax] += d
        return f(c)
    if len(axes) == 1:
        return (at([(axes[0], h)]) - at([(axes[0], -h)])) / (2 * h)
    if axes[0] == axes[1]:
        return (at([(axes[0], h)]) - 2 * f(np.array(centre, float)) + at([(axes[0], -h)])) / (h * h)
    return (at([(axes[0], h), (axes[1], h)]) - at([(axes[0], h), (axes[1], -h)]) - at([(axes[0], -h), (axes[1], h)])
            + at([(axes[0], -h), (axes[1], -h)])) / (4 * h * h)


A = np.array([0.1, -0.2, 0.3])
B = np.array([-0.4, 0.5, 0.9])
C = np.array([0.7, 0.2, -0.5])
NUC = np.array([[0.3, 0.4, -0.2]])
a, b, g = 0.9, 0.6, 1.3


def one_e(which, la, ca, centre_a):
    s, t, v = I.one_electron([prim(centre_a, la, a), prim(B, 1, b)], [2.0], NUC)
    m = {"S": s, "T": t, "V": v}[which]
    ia = I.cart_components(la).index(ca)
    return m[ia, (1 + la) * (2 + la) // 2 + 1] * raw_scale(la, a, ca) * raw_scale(1, b, (0, 1, 0))


def three_c(la, ca, centre_a, lc, cc, centre_c):
    raw = I.three_center([prim(centre_a, la, a), prim(B, 1, b)], [prim(centre_c, lc, g)])
    ia = I.cart_components(la).index(ca)
    ic = I.cart_components(lc).index(cc)
    nb = (1 + la) * (2 + la) // 2
    return raw[ic, ia, nb + 2] * raw_scale(la, a, ca) * raw_scale(1, b, (0, 0, 1)) * raw_scale(lc, g, cc)


@pytest.mark.parametrize("which", ["S", "T", "V"])
def test_one_electron_p_and_d_are_centre_derivatives(which):
    s_int = lambda c: one_e(which, 0, (0, 0, 0), c)
    for ax, comp in ((0, (1, 0, 0)), (1, (0, 1, 0)), (2, (0, 0, 1))):
        want = _fd(s_int, A, [ax]) / (2 * a)
        assert one_e(which, 1, comp, A) == pytest.approx(want, rel=1e-6, abs=1e-9)
    want = _fd(s_int, A, [0, 1]) / (4 * a * a)
    assert one_e(which, 2, (1, 1, 0), A) == pytest.approx(want, rel=1e-5, abs=1e-8)
    want = (_fd(s_int, A, [2, 2]) + 2 * a * s_int(A)) / (4 * a * a)
    assert one_e(which, 2, (0, 0, 2), A) == pytest.approx(want, rel=1e-5, abs=1e-8)


def test_three_centre_p_d_f_are_centre_derivatives():
    # orbital p and d on A, auxiliary p, d and f on C
    s_a = lambda c: three_c(0, (0, 0, 0), c, 0, (0, 0, 0), C)
    assert three_c(1, (0, 0, 1), A, 0, (0, 0, 0), C) == pytest.approx(_fd(s_a, A, [2]) / (2 * a), rel=1e-6)
    assert three_c(2, (1, 0, 1), A, 0, (0, 0, 0), C) == pytest.approx(_fd(s_a, A, [0, 2]) / (4 * a * a), rel=1e-5)
    s_c = lambda c: three_c(1, (1, 0, 0), A, 0, (0, 0, 0), c)
    assert three_c(1, (1, 0, 0), A, 1, (0, 1, 0), C) == pytest.approx(_fd(s_c, C, [1]) / (2 * g), rel=1e-6)
    want = (_fd(s_c, C, [0, 0]) + 2 * g * s_c(C)) / (4 * g * g)
    assert three_c(1, (1, 0, 0), A, 2, (2, 0, 0), C) == pytest.approx(want, rel=1e-5)
    # f_xyz = (1/2g) d/dC_z d_xy
    d_c = lambda c: three_c(1, (1, 0, 0), A, 2, (1, 1, 0), c)
    assert three_c(1, (1, 0, 0), A, 3, (1, 1, 1), C) == pytest.approx(_fd(d_c, C, [2]) / (2 * g), rel=1e-6)


def test_two_centre_p_and_d_are_centre_derivatives():
    def tc(lc, cc, cc_centre, ld, cd):
        m = I.two_center([prim(cc_centre, lc, g), prim(B, ld, b)])
        ic = I.cart_components(lc).index(cc)
        nd = (1 + lc) * (2 + lc) // 2
        return m[ic, nd + I.cart_components(ld).index(cd)] * raw_scale(lc, g, cc) * raw_scale(ld, b, cd)
    s_c = lambda c: tc(0, (0, 0, 0), c, 1, (0, 1, 0))
    assert tc(1, (1, 0, 0), C, 1, (0, 1, 0)) == pytest.approx(_fd(s_c, C, [0]) / (2 * g), rel=1e-6)
    assert tc(2, (0, 1, 1), C, 1, (0, 1, 0)) == pytest.approx(_fd(s_c, C, [1, 2]) / (4 * g * g), rel=1e-5)


def test_pauto_auxiliary_basis_recipe():
    basis = I.parse_gbs(open(os.path.join(GOLDEN, "lih-6-31g.gbs")).read())
    shells = I.molecule_shells([("Li", np.zeros(3)), ("H", np.array([0.0, 0.0, 3.0]))], basis)
    aux = I.auto_aux_shells(shells)
    for sym, lmax in (("Li", 2), ("H", 0)):  # lmax_abs = max(lmax_obs + 1, 2 lval(Z)), H capped by 2 lval
        mine = [s for s in aux if s.element == sym]
        assert max(s.l for s in mine) <= max(lmax, 0) + (1 if sym == "H" else 0)
        exps = [s.exps[0] for s in mine if s.l == 0]
        assert all(x > y for x, y in zip(exps, exps[1:]))  # clusters from the largest exponent down
        # every cluster emits l = 0 .. l_selected at one exponent, and l_selected never decreases
        ls = [s.l for s in mine]
        tops = [ls[i] for i in range(len(ls)) if i + 1 == len(ls) or ls[i + 1] == 0]
        assert all(x <= y for x, y in zip(tops, tops[1:]))
    # the largest Li cluster: the largest candidate 2 * 642.4189 merged with every candidate within 1.5
    li_s = sorted({float(p) for p in basis["Li"][0][1]}, reverse=True)
    assert aux[0].exps[0] == pytest.approx(2 * li_s[0])
    # AUTO: each primitive doubled only
    auto = I.auto_aux_shells(shells, "AUTO")
    assert len(auto) < len(aux)


def test_lih_ri_hf_energy_matches_the_reference():
    basis = I.parse_gbs(open(os.path.join(GOLDEN, "lih-6-31g.gbs")).read())
    mol = I.GaussianMolecule([("Li", np.zeros(3)), ("H", np.array([0.0, 0.0, 1.5949]) * I.BOHR_PER_ANGSTROM)], basis)
    ints = mol.integrals()
    assert mol.n == 11 and ints["n_aux_raw"] == 106 and ints["n_aux_dropped"] == 2
    s, h, b = ints["S"], ints["h"], ints["eri3"]
    w, u = np.linalg.eigh(s)
    x = u[:, w > 1e-6] / np.sqrt(w[w > 1e-6])
    fb = lambda p: oracle.nonrel_fock_ri(b.reshape(b.shape[0], -1), x, s, h, p)
    r = scf.nonrel_scf(fb, h, x, s, scf.core_guess(h, x, 4), 4, ints["e_nuc"], diis_size=5, energy_tolerance=1e-11,
                       density_tolerance=1e-9, max_iterations=100)
    assert r.converged
    assert abs(r.total_energy - (-7.9793032232)) < 1e-8


def _dense_hf(s, h, eri, ne):
    w, u = np.linalg.eigh(s)
    x = u[:, w > 1e-6] / np.sqrt(w[w > 1e-6])

    def fb(p):
        return h + np.einsum("mnls,sl->mn", eri, p) - 0.5 * np.einsum("mlsn,sl->mn", eri, p)

    r = scf.nonrel_scf(fb, h, x, s, scf.core_guess(h, x, ne), ne, 0.0, diis_size=5, energy_tolerance=1e-12,
                       density_tolerance=1e-10, max_iterations=200)
    assert r.converged
    return r.total_energy


def test_ne_dense_hf_energies_match_the_reference():
    basis = I.parse_gbs(open(os.path.join(GOLDEN, "ne-cc-pvdz.gbs")).read())
    shells = I.molecule_shells([("Ne", np.zeros(3))], basis)
    s, t, v = I.one_electron(shells, [10.0], [np.zeros(3)])
    nrm = I.normalization(np.diag(s))
    o2 = np.outer(nrm, nrm)
    s, h = s * o2, (t + v) * o2
    eri = I.four_center(shells) * np.einsum("a,b,c,d->abcd", nrm, nrm, nrm, nrm)
    assert abs(_dense_hf(s, h, eri, 10) - (-128.488866172)) < 1e-8
    # spherical d: the five real solid-harmonic combinations in place of the six Cartesian components
    n = s.shape[0]
    d0 = n - 6
    ix = {c: d0 + k for k, c in enumerate(I.cart_components(2))}
    tr = np.zeros((n, n - 1))
    tr[:d0, :d0] = np.eye(d0)
    cols = [{(1, 1, 0): 1}, {(1, 0, 1): 1}, {(0, 1, 1): 1}, {(2, 0, 0): 1, (0, 2, 0): -1},
            {(0, 0, 2): 2, (2, 0, 0): -1, (0, 2, 0): -1}]
    for k, col in enumerate(cols):
        for c, val in col.items():
            tr[ix[c], d0 + k] = val
    e_sph = _dense_hf(tr.T @ s @ tr, tr.T @ h @ tr, np.einsum("abcd,ai,bj,ck,dl->ijkl", eri, tr, tr, tr, tr), 10)
    assert abs(e_sph - (-128.4887755517)) < 1e-8
