"""The 4c RKB RI Dirac-Hartree-Fock SCF on the GPU Fock build (rkbFockMatrix, RI source), and the
learned starting density at that level.

- LiH/6-31G at the reference geometry reaches the reference's own 4c RI DHF energy, -7.9800761582 Ha
  (examples/reference.yaml:1239), and the energy of the same SCF run on the oracle's build.
- Learned guess: LiH with its bond length perturbed by a seeded U(-0.1, 0.1) Angstrom (the molecule
  stays on the z axis, so the densities share one frame); the network (1 distance -> 3 x 256 tanh ->
  the 44^2 real parameters of the RKB density) is trained on the device on the converged densities of
  the GPU SCF and must reach every held-out geometry's core-guess energy (1e-10 Ha) with fewer Fock
  builds."""
import multiprocessing
import os

import numpy as np
import pytest

from oracle import diis, oracle
from rerdmft_amd import dirac, guess, integrals as ig

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = dict(energy_tolerance=1e-11, density_tolerance=1e-9, max_iterations=80)


def lih(r_angstrom):
    return [("Li", np.zeros(3)), ("H", np.array([0.0, 0.0, r_angstrom]) * ig.BOHR_PER_ANGSTROM)]


def basis():
    return ig.parse_gbs(open(os.path.join(GOLDEN, "lih-6-31g.gbs")).read())


def test_lih_4c_ri_dhf_energy_on_the_gpu_build(ctx):
    mol = dirac.DiracMolecule(lih(1.5949), basis())
    res, ints = dirac.run_dhf(ctx, mol, slot_ll=12, slot_ss=13, **TOL)
    assert res.converged
    print(f"4c RI DHF LiH/6-31G: E = {res.total_energy:.10f} Ha in {res.iterations} Fock builds", flush=True)
    assert abs(res.total_energy - (-7.9800761582)) < 1e-8
    # the same loop on the oracle's build
    h, s, x, v = ints["h_rkb"], ints["s_full"], ints["x_full"], ints["v_total"]
    ell = ints["eri3_ll"].reshape(ints["eri3_ll"].shape[0], -1)
    ess = ints["eri3_ss"].reshape(ints["eri3_ss"].shape[0], -1)
    ro = dirac.dhf_scf(lambda p: oracle.rkb_fock_ri(ints["nl"], ints["ns"], ell, ess, v, s, x, h, p), h, x, s,
                       dirac.core_guess(h, x, 4), 4, ints["e_nuc"], diis=diis.Diis(8).extrapolate, **TOL)
    assert abs(res.total_energy - ro.total_energy) < 1e-10


def test_learned_guess_at_the_4c_level(ctx):
    g = np.random.default_rng(20261204)
    n_train, n_test = 64, 8
    lengths = 1.5949 + g.uniform(-0.1, 0.1, n_train + n_test)
    bs = basis()
    with multiprocessing.get_context("spawn").Pool(8) as pool:
        inputs = pool.map(dirac.geometry_inputs, [(lih(r), bs, {}) for r in lengths])
    positions = [np.array([a[1] for a in lih(r)]) for r in lengths]
    core = [dirac.run_dhf_inputs(ctx, ints, 4, slot_ll=12, slot_ss=13, **TOL) for ints in inputs]
    assert all(r.converged for r in core)
    net = guess.DhfDensityNet(ctx, n_atoms=2, n_rkb=inputs[0]["n_rkb"], n_electrons=4, seed=5)
    try:
        loss = net.fit(positions[:n_train], [r.density_matrix for r in core[:n_train]], steps=4000)
        builds_core, builds_mlp = [], []
        for k in range(n_train, n_train + n_test):
            ints = inputs[k]
            p0, source = net.initial_density(positions[k], ints["h_rkb"], ints["s_full"], ints["x_full"])
            assert source == "mlp"
            r = dirac.run_dhf_inputs(ctx, ints, 4, initial_density=p0, slot_ll=12, slot_ss=13, **TOL)
            assert r.converged
            assert abs(r.total_energy - core[k].total_energy) < 1e-10
            builds_core.append(core[k].iterations)
            builds_mlp.append(r.iterations)
    finally:
        net.close()
    print(f"4c learned guess (LiH, {n_train} training geometries, loss {loss:.2e}): Fock builds per held-out SCF "
          f"{np.mean(builds_core):.2f} (core guess) -> {np.mean(builds_mlp):.2f} (network)", flush=True)
    assert np.mean(builds_mlp) < np.mean(builds_core)
