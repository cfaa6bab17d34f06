"""CPU checks of the 4c RKB inputs (rerdmft_amd/dirac.py) and the DHF SCF loop, with the oracle's
rkbFockMatrix (RI source) as the Fock build:
- the exact Dirac hydrogen 1s energy c^2 (sqrt(1 - 1/c^2) - 1) from the one-electron RKB Hamiltonian
  in an even-tempered s basis (the small component, the sigma.p coupling and the -2c^2 shift), with
  Kramers-degenerate levels and the negative-energy branch below -2c^2;
- the reference's own 4c RI DHF energy of LiH/6-31G, -7.9800761582 Ha (examples/reference.yaml:1239,
  lih_gnof_c4_ri_4c; basis tests/golden/lih-6-31g.gbs, the reference's examples/lih-6-31g.gbs);
- the non-relativistic limit: E(c) - E_NR = -k / c^2 (E_NR the NON_REL RI HF energy with the same
  auxiliary basis and the large-component tensor), checked at c = 1e3 and 1e4."""
import os

import numpy as np
import pytest

from oracle import diis, oracle
from rerdmft_amd import dirac, integrals as ig, scf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LIH = [("Li", np.zeros(3)), ("H", np.array([0.0, 0.0, 1.5949]) * ig.BOHR_PER_ANGSTROM)]


def lih_basis():
    return ig.parse_gbs(open(os.path.join(GOLDEN, "lih-6-31g.gbs")).read())


def oracle_dhf(mol, tol=1e-12):
    ints = mol.integrals()
    h, s, x, v = ints["h_rkb"], ints["s_full"], ints["x_full"], ints["v_total"]
    ell = ints["eri3_ll"].reshape(ints["eri3_ll"].shape[0], -1)
    ess = ints["eri3_ss"].reshape(ints["eri3_ss"].shape[0], -1)
    fb = lambda p: oracle.rkb_fock_ri(ints["nl"], ints["ns"], ell, ess, v, s, x, h, p)  # noqa: E731
    r = dirac.dhf_scf(fb, h, x, s, dirac.core_guess(h, x, mol.n_electrons), mol.n_electrons, ints["e_nuc"],
                      diis=diis.Diis(8).extrapolate, max_iterations=80, energy_tolerance=tol, density_tolerance=tol * 100)
    assert r.converged
    return r, ints


def test_hydrogen_like_dirac_1s_energy():
    basis = {"H": [(0, np.array([a]), np.array([1.0])) for a in 0.01 * 2.0 ** np.arange(40)]}
    mol = dirac.DiracMolecule([("H", np.zeros(3))], basis, n_electrons=1)
    ints = mol.integrals()
    w, _ = dirac._spinors(ints["h_rkb"], ints["x_full"])
    c = dirac.SPEED_OF_LIGHT
    n = len(w)
    exact = c * c * (np.sqrt(1.0 - 1.0 / c ** 2) - 1.0)
    assert abs(w[n // 2] - exact) < 1e-8          # -0.50000665660 (non-relativistic -0.5)
    assert abs(w[n // 2 + 1] - w[n // 2]) < 1e-7  # Kramers pair (to the conditioning of this basis)
    assert w[n // 2 - 1] < -c * c                  # the negative-energy branch lies near -2c^2


def test_lih_4c_ri_dhf_energy_matches_the_reference():
    mol = dirac.DiracMolecule(LIH, lih_basis())
    r, ints = oracle_dhf(mol)
    assert ints["nl"] == 11 and ints["ns"] == 39 and ints["n_rkb"] == 44
    assert abs(r.total_energy - (-7.9800761582)) < 1e-8


def test_lih_4c_non_relativistic_limit():
    basis = lih_basis()
    gi = ig.GaussianMolecule(LIH, basis).integrals()
    s, h = gi["S"], gi["h"]
    w, u = np.linalg.eigh(s)
    x = u / np.sqrt(w)
    de = []
    for c in (1e3, 1e4):
        r, ints = oracle_dhf(dirac.DiracMolecule(LIH, basis, speed_of_light=c))
        b = ints["eri3_ll"].reshape(ints["eri3_ll"].shape[0], -1)
        rn = scf.nonrel_scf(lambda p: oracle.nonrel_fock_ri(b, x, s, h, p), h, x, s, scf.core_guess(h, x, 4), 4,
                            gi["e_nuc"], diis_size=5, energy_tolerance=1e-12, density_tolerance=1e-10)
        de.append(r.total_energy - rn.total_energy)
    assert de[0] < 0 and de[1] < 0
    assert de[0] / de[1] == pytest.approx(100.0, rel=2e-3)
    assert abs(de[1]) < 2e-7
