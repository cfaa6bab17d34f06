"""The learned SCF starting density on the GPU: the device network's forward pass, loss gradient and
Adam step against the FP64 oracle (oracle/mlp.py, itself pinned by finite differences), then the
whole loop -- converged densities of the GPU SCF on seeded perturbed water geometries, training on
the device, and held-out geometries started from the network's guess -- which must reach the same
FP64 energy as the core-guess run (1e-10 Ha) with fewer Fock (J/K) builds."""
import numpy as np
import pytest

from oracle import mlp as om
from rerdmft_amd import guess, molecule, scf

pytestmark = pytest.mark.gpu

SLOT = 15
DIMS = om.layer_dims(3, 256, 3, 210)  # water: 3 distances -> 3 x 256 -> 20*21/2 density entries
TOL = dict(diis_size=5, energy_tolerance=1e-11, density_tolerance=1e-8, max_iterations=60)


@pytest.fixture(scope="module")
def net(ctx):
    h = ctx.mlp_create(DIMS[0], 256, 3, DIMS[-1], seed=11)
    yield h
    ctx.mlp_destroy(h)


def _data(seed, batch=37):
    g = np.random.default_rng(seed)
    return g.standard_normal((batch, DIMS[0])), 0.5 * g.standard_normal((batch, DIMS[-1]))


def test_parameter_count_and_init(ctx, net):
    p = ctx.mlp_get_params(net)
    assert p.size == sum(fi * fo + fo for fi, fo in zip(DIMS[:-1], DIMS[1:]))
    layers = om.unpack(p, DIMS)
    for (w, b), fi, fo in zip(layers, DIMS[:-1], DIMS[1:]):
        lim = np.sqrt(6.0 / (fi + fo))
        assert np.all(np.abs(w) <= lim) and np.std(w) > 0.4 * lim
        assert np.all(b == 0)


def test_forward_matches_oracle(ctx, net):
    x, _ = _data(1)
    p = ctx.mlp_get_params(net).astype(np.float64)
    y = ctx.mlp_forward(net, x, DIMS[-1])
    ref = om.forward(p, DIMS, x)
    assert np.max(np.abs(y - ref)) <= 2e-5 * np.max(np.abs(ref))


def test_loss_gradient_matches_oracle(ctx, net):
    x, t = _data(2)
    p = ctx.mlp_get_params(net).astype(np.float64)
    loss, g = ctx.mlp_loss_grad(net, x, t)
    ref_loss, ref_g = om.loss_grad(p, DIMS, x, t)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    # per layer: FP32 products over K <= 256 terms
    off = 0
    for fi, fo in zip(DIMS[:-1], DIMS[1:]):
        for cnt in (fi * fo, fo):
            a, b = g[off:off + cnt], ref_g[off:off + cnt]
            assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(b)) + 1e-9
            off += cnt
    assert off == g.size


def test_adam_steps_match_oracle(ctx):
    dims = om.layer_dims(5, 64, 2, 9)
    h = ctx.mlp_create(5, 64, 2, 9, seed=5)
    try:
        g = np.random.default_rng(3)
        x, t = g.standard_normal((16, 5)), g.standard_normal((16, 9))
        p = ctx.mlp_get_params(h).astype(np.float64)
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        for step in (1, 2, 3):
            loss, grad = ctx.mlp_loss_grad(h, x, t)
            l_dev = ctx.mlp_train_step(h, x, t, 1e-3)
            assert l_dev == pytest.approx(loss, rel=1e-6)
            p_ref, m, v = om.adam_step(p, grad.astype(np.float64), m, v, step, 1e-3)
            p = ctx.mlp_get_params(h).astype(np.float64)
            # the update is lr * m_hat / sqrt(v_hat): compare where the gradient is not at round-off level
            big = np.abs(grad) > 1e-3 * np.max(np.abs(grad))
            assert np.max(np.abs(p - p_ref)[big]) <= 2e-6
        # training lowers the loss
        before = ctx.mlp_loss_grad(h, x, t)[0]
        for _ in range(200):
            ctx.mlp_train_step(h, x, t, 3e-3)
        assert ctx.mlp_loss_grad(h, x, t)[0] < 0.5 * before
    finally:
        ctx.mlp_destroy(h)


def test_rejects_bad_arguments(ctx, net):
    import rerdmft_amd as rr
    with pytest.raises(rr.GpuError):
        ctx.mlp_create(0, 8, 1, 1)
    with pytest.raises(ValueError):
        ctx.mlp_set_params(net, np.zeros(3, dtype=np.float32))
    with pytest.raises(rr.GpuError):
        ctx.mlp_loss_grad(net, np.zeros((0, 3)), np.zeros((0, 210)))


def test_gpu_scf_matches_oracle_scf(ctx, oracle):
    mol = molecule.water(molecule.perturbed_geometries(molecule.WATER_GEOMETRY, 1, 0.1, 99)[0])
    res, ints, src = guess.run_scf(ctx, mol, slot=SLOT, **TOL)
    assert res.converged and src == "core"
    s, h, x = ints["S"], ints["h"], ints["X"]
    b = ints["eri3"].reshape(mol.n_aux, -1)
    ref = scf.nonrel_scf(lambda p: oracle.nonrel_fock_ri(b, x, s, h, p), h, x, s, scf.core_guess(h, x, 10), 10,
                         ints["e_nuc"], **TOL)
    assert ref.converged
    assert abs(res.total_energy - ref.total_energy) < 1e-10
    assert res.iterations == ref.iterations


def test_trained_guess_cuts_fock_builds(ctx):
    train = molecule.perturbed_geometries(molecule.WATER_GEOMETRY, 128, 0.1, 20261203)
    held_out = molecule.perturbed_geometries(molecule.WATER_GEOMETRY, 8, 0.1, 20261204)
    dens, core_builds = [], 0
    for pos in train:
        r, _, _ = guess.run_scf(ctx, molecule.water(pos), slot=SLOT, **TOL)
        assert r.converged
        dens.append(r.density_matrix)
        core_builds += r.iterations
    mol0 = molecule.water()
    g = guess.DensityGuessNet(ctx, 3, mol0.n, mol0.n_electrons, seed=7)
    try:
        final = g.fit(train, dens, steps=10000)
        assert final < 1e-3 * g.losses[0]
        rows = []
        for pos in held_out:
            mol = molecule.water(pos)
            rc, _, _ = guess.run_scf(ctx, mol, slot=SLOT, **TOL)
            rm, _, src = guess.run_scf(ctx, mol, guess=g, slot=SLOT, **TOL)
            assert rc.converged and rm.converged and src == "mlp"
            assert abs(rm.total_energy - rc.total_energy) < 1e-10
            rows.append((rc.iterations, rm.iterations))
        rows = np.array(rows)
        print("held-out Fock builds (core guess, network guess):", rows.tolist(),
              "mean", rows.mean(axis=0).tolist(), "training-set mean (core):", core_builds / len(train))
        assert np.all(rows[:, 1] <= rows[:, 0])
        assert rows[:, 1].mean() <= rows[:, 0].mean() - 1.5
    finally:
        g.close()
