"""Molecule-level parity through the GPU RI Fock build: LiH/6-31G with the PAUTO auxiliary basis
(integrals from rerdmft_amd.integrals), RI NON_REL HF through rr_nonrel_fock_ri (nonRelFockMatrix's
RI path inside the reference's SCF loop, NonRelHartreeFock.cpp:118-198) must reproduce the reference's
own energy, examples/reference.yaml:990, -7.9793032232 Ha, within 1e-8 Ha; so must the
mixed-precision SCF (BF16 incremental builds, FP64 final iterations)."""
import os

import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import guess, integrals, scf

pytestmark = pytest.mark.gpu

E_REF = -7.9793032232
TOL = dict(diis_size=5, energy_tolerance=1e-11, density_tolerance=1e-9, max_iterations=100)


@pytest.fixture(scope="module")
def lih():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lih-6-31g.gbs")
    basis = integrals.parse_gbs(open(path).read())
    return integrals.GaussianMolecule([("Li", np.zeros(3)),
                                       ("H", np.array([0.0, 0.0, 1.5949]) * integrals.BOHR_PER_ANGSTROM)], basis)


def test_lih_ri_hf_on_the_gpu_matches_the_reference(ctx, lih):
    print(f"LiH/6-31G: n={lih.n} n_aux={lih.n_aux}", flush=True)
    r, ints, _ = guess.run_scf(ctx, lih, slot=13, copies=rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16, **TOL)
    print(f"  RI HF {r.total_energy:.10f} ({r.iterations} builds), reference {E_REF}", flush=True)
    assert r.converged and abs(r.total_energy - E_REF) < 1e-8
    full, delta = scf.gpu_builds(ctx, 13, rr.BF16)
    m = scf.nonrel_scf_incremental(full, delta, ints["h"], ints["X"], ints["S"], lih.n_electrons, ints["e_nuc"], **TOL)
    print(f"  mixed: {m.full_builds} FP64 + {m.delta_builds} BF16 builds, {m.total_energy:.10f}", flush=True)
    assert m.converged and abs(m.total_energy - E_REF) < 1e-8
    ctx.release(13)
