"""Learned SCF starting density: a small FP32 network on the GPU maps a geometry to the converged
density of the SCF method itself, replacing the core-Hamiltonian guess (main.cpp:5240-5250) when its
prediction is usable and falling back to that guess otherwise.

The network (rr_mlp_* in include/rerdmft_gpu.h: 3 hidden tanh layers of width 256 by default) is
trained on the device with its own forward/backward kernels and Adam. Inputs are the interatomic
distances of a fixed atom ordering (bohr; the AO density of an s-type basis is invariant under
rigid motions, so these determine it), standardised with the training statistics; targets are the
upper triangle of the converged AO density, centred on the training mean and scaled by its overall
standard deviation.

A prediction is turned into a valid closed-shell density before it is used: the eigenvectors of
X^T S P S X / 2 with the n_electrons/2 largest eigenvalues give orbitals C = X V, and P0 = 2 C C^T.
It is rejected (core guess instead) when it is not finite or when the occupied and virtual
eigenvalues are not separated by 1/2 (the occupation cut of NonRelFockMatrixRi.cpp:40-72).
"""
import numpy as np

from . import scf as _scf


def distances(positions):
    """Interatomic distances r_ij, i > j, in a fixed order."""
    r = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    return np.array([np.linalg.norm(r[i] - r[j]) for i in range(len(r)) for j in range(i)])


def pack_triangle(p):
    iu = np.triu_indices(p.shape[0])
    return p[iu]


def unpack_triangle(v, n):
    p = np.zeros((n, n))
    iu = np.triu_indices(n)
    p[iu] = v
    return p + np.triu(p, 1).T


def purify(p, overlap, x_large, n_electrons):
    """Closed-shell density from an approximate one, or None when the occupied set is ambiguous."""
    if not np.all(np.isfinite(p)):
        return None
    no = n_electrons // 2
    s = overlap
    m = x_large.T @ (s @ p @ s) @ x_large * 0.5
    w, v = np.linalg.eigh(0.5 * (m + m.T))
    w, v = w[::-1], v[:, ::-1]
    if no == 0 or w[no - 1] <= 0.5 or (len(w) > no and w[no] >= 0.5):
        return None
    c = x_large @ v[:, :no]
    return 2.0 * c @ c.T


class DensityGuessNet:
    """Geometry -> converged density network for one molecule family (fixed atoms and basis)."""

    def __init__(self, ctx, n_atoms, n_basis, n_electrons, hidden=256, layers=3, seed=1):
        self.ctx = ctx
        self.n = n_basis
        self.n_electrons = n_electrons
        self.n_in = n_atoms * (n_atoms - 1) // 2
        self.n_out = n_basis * (n_basis + 1) // 2
        self.handle = ctx.mlp_create(self.n_in, hidden, layers, self.n_out, seed)
        self.x_mean = self.x_std = self.t_mean = None
        self.t_scale = 1.0
        self.losses = []

    def close(self):
        if self.handle is not None:
            self.ctx.mlp_destroy(self.handle)
            self.handle = None

    def _features(self, positions):
        x = np.array([distances(p) for p in positions])
        return (x - self.x_mean) / self.x_std

    def _pack(self, p):
        return pack_triangle(p)

    def _unpack(self, v):
        return unpack_triangle(v, self.n)

    def fit(self, positions, densities, steps=10000, lr=1e-3, lr_final=1e-5):
        """Full-batch Adam on the GPU; the learning rate decays geometrically from lr to lr_final."""
        x = np.array([distances(p) for p in positions])
        t = np.array([self._pack(d) for d in densities])
        self.x_mean = x.mean(axis=0)
        self.x_std = np.maximum(x.std(axis=0), 1e-6)
        self.t_mean = t.mean(axis=0)
        self.t_scale = max(float((t - self.t_mean).std()), 1e-12)
        xs = (x - self.x_mean) / self.x_std
        ts = (t - self.t_mean) / self.t_scale
        decay = (lr_final / lr) ** (1.0 / max(steps - 1, 1))
        rate = lr
        for _ in range(steps):
            self.losses.append(self.ctx.mlp_train_step(self.handle, xs, ts, rate))
            rate *= decay
        return self.losses[-1]

    def predict_raw(self, positions):
        y = self.ctx.mlp_forward(self.handle, self._features(positions), self.n_out)
        return [self._unpack(self.t_mean + self.t_scale * row) for row in y]

    def initial_density(self, positions, h_core, overlap, x_large):
        """(P0, source): the purified prediction ("mlp") or the core-Hamiltonian guess ("core")."""
        p = purify(self.predict_raw([positions])[0], overlap, x_large, self.n_electrons)
        if p is None:
            return _scf.core_guess(h_core, x_large, self.n_electrons), "core"
        return p, "mlp"


def pack_hermitian(p):
    """Real parameter vector of a Hermitian matrix: Re of the upper triangle, then Im above the diagonal."""
    n = p.shape[0]
    return np.concatenate([p.real[np.triu_indices(n)], p.imag[np.triu_indices(n, 1)]])


def unpack_hermitian(v, n):
    nt = n * (n + 1) // 2
    re, im = np.zeros((n, n)), np.zeros((n, n))
    re[np.triu_indices(n)] = v[:nt]
    im[np.triu_indices(n, 1)] = v[nt:]
    return (re + np.triu(re, 1).T) + 1j * (im - im.T)


def purify_rkb(p, s_full, x_full, n_electrons):
    """RKB density from an approximate one: the eigenvectors of X^H S P S X with the n_electrons
    largest eigenvalues give spinors C = X V and P0 = C C^H; None when not finite or when the occupied
    and the other eigenvalues are not separated by 1/2."""
    if not np.all(np.isfinite(p)):
        return None
    m = x_full.conj().T @ (s_full @ p @ s_full) @ x_full
    w, v = np.linalg.eigh(0.5 * (m + m.conj().T))
    w, v = w[::-1], v[:, ::-1]
    ne = n_electrons
    if ne == 0 or w[ne - 1] <= 0.5 or (len(w) > ne and w[ne] >= 0.5):
        return None
    c = x_full @ v[:, :ne]
    return c @ c.conj().T


class DhfDensityNet(DensityGuessNet):
    """Geometry -> converged 4c RKB density (complex Hermitian, n_rkb x n_rkb) for one molecule family;
    the fallback is the 4c core guess (main.cpp:6117-6119). Targets are the n_rkb^2 real parameters of
    pack_hermitian."""

    def __init__(self, ctx, n_atoms, n_rkb, n_electrons, hidden=256, layers=3, seed=1):
        self.ctx = ctx
        self.n = n_rkb
        self.n_electrons = n_electrons
        self.n_in = n_atoms * (n_atoms - 1) // 2
        self.n_out = n_rkb * n_rkb
        self.handle = ctx.mlp_create(self.n_in, hidden, layers, self.n_out, seed)
        self.x_mean = self.x_std = self.t_mean = None
        self.t_scale = 1.0
        self.losses = []

    def _pack(self, p):
        return pack_hermitian(p)

    def _unpack(self, v):
        return unpack_hermitian(v, self.n)

    def initial_density(self, positions, h_rkb, s_full, x_full):
        """(P0, source): the purified prediction ("mlp") or the 4c core guess ("core")."""
        from . import dirac
        p = purify_rkb(self.predict_raw([positions])[0], s_full, x_full, self.n_electrons)
        if p is None:
            return dirac.core_guess(h_rkb, x_full, self.n_electrons), "core"
        return p, "mlp"


def run_scf(ctx, mol, guess=None, slot=0, lin_dep=1e-10, copies=None, **scf_options):
    """NON_REL RI SCF of a Molecule with the GPU Fock build (nonRelFockMatrix RI path).

    guess: None for the core-Hamiltonian starting density, or a trained DensityGuessNet (which
    itself falls back to the core guess). copies: the tensor copies to keep (default FP64 only). Returns (ScfResult, ints, source) with ints the integral
    dict of the molecule plus its orthogonaliser X, and source "core" or "mlp"."""
    ints = mol.integrals()
    s, h = ints["S"], ints["h"]
    x = ctx.canonical_orthogonalize(s, lin_dep)
    ints["X"] = x
    ctx.upload(slot, ints["eri3"]) if copies is None else ctx.upload(slot, ints["eri3"], copies)
    if guess is None:
        p0, source = _scf.core_guess(h, x, mol.n_electrons), "core"
    else:
        p0, source = guess.initial_density(mol.positions, h, s, x)
    res = _scf.nonrel_scf(lambda p: ctx.nonrel_fock_ri(slot, h, x, s, p), h, x, s, p0, mol.n_electrons,
                          ints["e_nuc"], **scf_options)
    return res, ints, source
