"""The 4c RKB RI Dirac-Hartree-Fock SCF on the GPU Fock build (rkbFockMatrix, RI source), and the
learned starting density at that level.

- LiH/6-31G at the reference geometry reaches the reference's own 4c RI DHF energy, -7.9800761582 Ha
  (examples/reference.yaml:1239), and the energy of the same SCF run on the oracle's build.
- Learned guess: LiH/6-31G with its bond length perturbed by a seeded U(-0.1, 0.1) Angstrom (the
  molecule stays on the z axis, so the densities share one frame), and water/STO-3G with every
  coordinate perturbed by U(-0.1, 0.1) Angstrom and the result placed in a canonical frame; the network
  (interatomic distances -> 3 x 256 tanh -> the n_rkb^2 real parameters of the RKB density) is trained
  on the device on the converged densities of the GPU SCF and must reach every held-out geometry's
  core-guess energy (1e-10 Ha) with fewer Fock builds."""
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


def water(r1, r2, theta):
    """Water in a canonical frame fixed by its internal coordinates (bond lengths in Angstrom, angle in
    radians): O at the origin, the HOH bisector along +z, the molecule in the yz plane. With p functions
    the AO density depends on the orientation, so every geometry is placed in this frame."""
    a = ig.BOHR_PER_ANGSTROM
    h1 = r1 * np.array([0.0, np.sin(theta / 2), np.cos(theta / 2)]) * a
    h2 = r2 * np.array([0.0, -np.sin(theta / 2), np.cos(theta / 2)]) * a
    return [("O", np.zeros(3)), ("H", h1), ("H", h2)]


def perturbed_water(g):
    """Each atom of the reference geometry (O at the origin, H at (0, +-0.757, 0.586) Angstrom) moved by a
    seeded U(-0.1, 0.1) Angstrom per coordinate, then put in the canonical frame."""
    ref = np.array([[0.0, 0.0, 0.0], [0.0, 0.757, 0.586], [0.0, -0.757, 0.586]])
    x = ref + g.uniform(-0.1, 0.1, ref.shape)
    d1, d2 = x[1] - x[0], x[2] - x[0]
    r1, r2 = np.linalg.norm(d1), np.linalg.norm(d2)
    return water(r1, r2, np.arccos(np.dot(d1, d2) / (r1 * r2)))


def learned_guess(ctx, geometries, bs, n_electrons, n_train, seed, steps):
    """Host inputs in worker processes, core-guess SCFs on the GPU build, device training on the first
    n_train converged densities, then every other geometry from the network's guess: returns the Fock
    builds per held-out SCF (core guess, network), the final loss and the largest energy difference."""
    with multiprocessing.get_context("spawn").Pool(min(16, os.cpu_count() or 1)) as pool:
        inputs = pool.map(dirac.geometry_inputs, [(geo, bs, {}) for geo in geometries])
    positions = [np.array([a[1] for a in geo]) for geo in geometries]
    core = [dirac.run_dhf_inputs(ctx, ints, n_electrons, slot_ll=12, slot_ss=13, **TOL) for ints in inputs]
    assert all(r.converged for r in core)
    net = guess.DhfDensityNet(ctx, n_atoms=len(geometries[0]), n_rkb=inputs[0]["n_rkb"], n_electrons=n_electrons,
                              seed=seed)
    builds_core, builds_mlp, de = [], [], 0.0
    try:
        loss = net.fit(positions[:n_train], [r.density_matrix for r in core[:n_train]], steps=steps)
        for k in range(n_train, len(geometries)):
            ints = inputs[k]
            p0, source = net.initial_density(positions[k], ints["h_rkb"], ints["s_full"], ints["x_full"])
            assert source == "mlp"
            r = dirac.run_dhf_inputs(ctx, ints, n_electrons, initial_density=p0, slot_ll=12, slot_ss=13, **TOL)
            assert r.converged
            de = max(de, abs(r.total_energy - core[k].total_energy))
            builds_core.append(core[k].iterations)
            builds_mlp.append(r.iterations)
    finally:
        net.close()
    return np.mean(builds_core), np.mean(builds_mlp), loss, de


def test_learned_guess_at_the_4c_level_lih(ctx):
    g = np.random.default_rng(20261204)
    n_train, n_test = 256, 16
    geos = [lih(r) for r in 1.5949 + g.uniform(-0.1, 0.1, n_train + n_test)]
    core, mlp, loss, de = learned_guess(ctx, geos, basis(), 4, n_train, 5, 4000)
    print(f"4c learned guess (LiH/6-31G, {n_train} training geometries, loss {loss:.2e}): Fock builds per held-out "
          f"SCF {core:.2f} (core guess) -> {mlp:.2f} (network), max |dE| {de:.1e} Ha", flush=True)
    assert de < 1e-10
    assert mlp < core


def test_learned_guess_at_the_4c_level_water(ctx):
    g = np.random.default_rng(20261205)
    n_train, n_test = 384, 16
    geos = [perturbed_water(g) for _ in range(n_train + n_test)]
    bs = ig.parse_gbs(open(os.path.join(GOLDEN, "water-sto-3g.gbs")).read())
    core, mlp, loss, de = learned_guess(ctx, geos, bs, 10, n_train, 6, 12000)
    print(f"4c learned guess (water/STO-3G, {n_train} training geometries, loss {loss:.2e}): Fock builds per "
          f"held-out SCF {core:.2f} (core guess) -> {mlp:.2f} (network), max |dE| {de:.1e} Ha", flush=True)
    assert de < 1e-10
    assert mlp < core
