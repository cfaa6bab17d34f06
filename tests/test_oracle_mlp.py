"""CPU checks of the learned-guess stack: the FP64 network oracle (finite differences, a planted
mistake, Adam's first step), the s-type integral engine (published H2/STO-3G values), the RI fit,
the SCF loop and the density purification."""
import numpy as np
import pytest

from oracle import mlp as om
from oracle import oracle
from rerdmft_amd import guess, molecule, scf


def _net(seed, dims):
    g = np.random.default_rng(seed)
    parts = []
    for fi, fo in zip(dims[:-1], dims[1:]):
        parts += [g.uniform(-1, 1, fi * fo) / np.sqrt(fi), 0.1 * g.standard_normal(fo)]
    return np.concatenate(parts)


@pytest.mark.parametrize("planted", [False, True])
def test_oracle_gradient_matches_central_differences(planted):
    dims = om.layer_dims(3, 16, 3, 5)
    p = _net(1, dims)
    g = np.random.default_rng(2)
    x, t = g.standard_normal((7, 3)), g.standard_normal((7, 5))
    _, grad = om.loss_grad(p, dims, x, t, planted=planted)
    h = 1e-5
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (om.loss(p + e, dims, x, t) - om.loss(p - e, dims, x, t)) / (2 * h)
    err = np.max(np.abs(grad - fd)) / np.max(np.abs(fd))
    if planted:
        assert err > 1e-3  # the planted derivative mistake is caught
    else:
        assert err < 1e-7


def test_oracle_loss_value():
    dims = om.layer_dims(2, 4, 1, 3)
    p = _net(3, dims)
    x, t = np.ones((2, 2)), np.zeros((2, 3))
    y = om.forward(p, dims, x)
    assert om.loss(p, dims, x, t) == pytest.approx(0.5 * np.sum(y * y) / 6, rel=1e-14)
    # one hidden layer by hand
    (w1, b1), (w2, b2) = om.unpack(p, dims)
    np.testing.assert_allclose(y, np.tanh(x @ w1.T + b1) @ w2.T + b2, rtol=1e-14)


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([3.0, -1e-3, 0.5, -7.0])
    p, m, v = om.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def _h2():
    z = 1.24  # Szabo and Ostlund, section 3.5.2: STO-3G H2 at R = 1.4 bohr, zeta = 1.24
    spec = {1: [(molecule.STO3G_1S_EXP * z * z, molecule.STO3G_1S_COEF)]}
    return molecule.Molecule([1, 1], [[0, 0, 0], [1.4, 0, 0]], spec, {1: molecule.even_tempered(0.2, 2.0, 10)})


def test_integrals_match_szabo_ostlund_h2():
    mol = _h2()
    s, t, v = molecule.one_electron(mol.basis, mol.charges, mol.positions)
    _, _, v1 = molecule.one_electron(mol.basis, [1.0], [[0, 0, 0]])
    g = molecule.four_center(mol.basis)
    pub = [(s[0, 1], 0.6593), (t[0, 0], 0.7600), (t[0, 1], 0.2365), (v1[0, 0], -1.2266), (v1[0, 1], -0.5974),
           (v1[1, 1], -0.6538), (g[0, 0, 0, 0], 0.7746), (g[0, 0, 1, 1], 0.5697), (g[1, 0, 0, 0], 0.4441),
           (g[1, 0, 1, 0], 0.2970)]
    for got, want in pub:
        assert abs(got - want) < 6e-5, (got, want)
    assert abs(s[0, 0] - 1.0) < 1e-14


def test_scf_loop_reproduces_h2_hartree_fock_energy():
    mol = _h2()
    s, t, v = molecule.one_electron(mol.basis, mol.charges, mol.positions)
    g = molecule.four_center(mol.basis)
    h = t + v

    def fock(p):
        return h + np.einsum("abcd,dc->ab", g, p) - 0.5 * np.einsum("adcb,dc->ab", g, p)

    x = molecule.inverse_sqrt(s)
    r = scf.nonrel_scf(fock, h, x, s, scf.core_guess(h, x, 2), 2, molecule.nuclear_repulsion(mol.charges, mol.positions),
                       energy_tolerance=1e-13, density_tolerance=1e-11)
    assert r.converged
    assert r.electronic_energy == pytest.approx(-1.8310, abs=1e-4)
    assert r.total_energy == pytest.approx(-1.1167, abs=1e-4)


def test_ri_fit_is_exact_when_aux_spans_products():
    # one centre: products of s primitives are s functions with the summed exponent
    exps = [0.5, 2.0]
    spec = {1: [([e], [1.0]) for e in exps]}
    prods = sorted({a + b for a in exps for b in exps})
    mol = molecule.Molecule([1], [[0.1, -0.2, 0.3]], spec, {1: np.array(prods)}, n_electrons=2)
    b = mol.integrals()["eri3"]
    g = molecule.four_center(mol.basis)
    np.testing.assert_allclose(np.einsum("pab,pcd->abcd", b, b), g, atol=1e-10)


def test_water_energy_is_invariant_under_rigid_motion():
    def energy(pos):
        mol = molecule.water(pos)
        ints = mol.integrals()
        s, h = ints["S"], ints["h"]
        x = oracle.canonical_orthogonalize(s, 1e-10)
        b = ints["eri3"].reshape(mol.n_aux, -1)
        r = scf.nonrel_scf(lambda p: oracle.nonrel_fock_ri(b, x, s, h, p), h, x, s, scf.core_guess(h, x, 10), 10,
                           ints["e_nuc"], energy_tolerance=1e-12, density_tolerance=1e-9)
        assert r.converged
        return r.total_energy, r.density_matrix

    e0, p0 = energy(molecule.WATER_GEOMETRY)
    c, sn = np.cos(0.7), np.sin(0.7)
    rot = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1]])
    e1, p1 = energy(molecule.WATER_GEOMETRY @ rot.T + np.array([0.4, -0.3, 1.1]))
    assert abs(e1 - e0) < 1e-10
    np.testing.assert_allclose(p1, p0, atol=1e-7)
    assert -80 < e0 < -60


def test_purify_gives_closed_shell_density():
    mol = molecule.water()
    ints = mol.integrals()
    s, h = ints["S"], ints["h"]
    x = oracle.canonical_orthogonalize(s, 1e-10)
    p0 = scf.core_guess(h, x, 10)
    g = np.random.default_rng(4).standard_normal(p0.shape)
    p = guess.purify(p0 + 1e-3 * (g + g.T), s, x, 10)
    assert p is not None
    np.testing.assert_allclose(p @ s @ p, 2.0 * p, atol=1e-9)
    assert np.trace(p @ s) == pytest.approx(10.0, abs=1e-10)
    np.testing.assert_allclose(guess.purify(p0, s, x, 10), p0, atol=1e-9)
    assert guess.purify(np.full_like(p0, np.nan), s, x, 10) is None
    assert guess.purify(np.zeros_like(p0), s, x, 10) is None  # no occupied set


def test_triangle_packing_round_trip():
    a = np.random.default_rng(5).standard_normal((6, 6))
    a = a + a.T
    v = guess.pack_triangle(a)
    assert v.size == 21
    np.testing.assert_array_equal(guess.unpack_triangle(v, 6), a)


def test_distances_invariant():
    pos = molecule.WATER_GEOMETRY
    d = guess.distances(pos)
    assert d.shape == (3,)
    assert d[0] == pytest.approx(np.linalg.norm(pos[1] - pos[0]))
    np.testing.assert_allclose(guess.distances(pos + 3.0), d)
