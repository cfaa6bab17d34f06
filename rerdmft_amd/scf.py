"""The NON_REL restricted Hartree-Fock SCF loop around the device Fock build.

Mirrors nonRelScfImpl (src/NON_REL/NonRelHartreeFock.cpp:118-198): per iteration one Fock build
F(P) (here `fock_build`, normally Context.nonrel_fock_ri -- the RI J/K build on the GPU), Pulay DIIS
on the commutator error F P S - S P F (src/Utils/DIIS.h:10-38, DIIS.cpp:79-130), diagonalisation
of X^T F X, the new density from the lowest n_electrons/2 orbitals (nonRelDensityMatrix,
NonRelHartreeFock.cpp:42-70), the energy 1/2 tr(P (h + F)) with the density that built F, and
convergence as soon as either the density change or the energy change is below its tolerance
(OR, from the second iteration on). The small dense FP64 steps (DIIS solve, eigensolve) run on the
host, like the reference's.

The core-Hamiltonian guess is main.cpp:5240-5250: P0 from the eigenvectors of X^T h X.
"""
from collections import deque
from dataclasses import dataclass, field

import numpy as np


def density_matrix(c, n_electrons):
    """nonRelDensityMatrix: 2 sum_{i < n_e/2} c_i c_i^T (NonRelHartreeFock.cpp:42-70)."""
    n = c.shape[0]
    if c.shape[1] != n:
        raise ValueError("nonRelDensityMatrix: C is not square")
    if n_electrons <= 0 or n_electrons % 2:
        raise ValueError("nonRelDensityMatrix: n_electrons must be a positive even number")
    no = n_electrons // 2
    if no > n:
        raise ValueError("nonRelDensityMatrix: n_electrons/2 exceeds the number of orbitals")
    co = c[:, :no]
    return 2.0 * co @ co.T


def _orbitals(f, x):
    w, v = np.linalg.eigh(x.T @ f @ x)
    c = x @ v
    if c.shape[1] < c.shape[0]:  # rectangular X: pad with zero columns as the square C of the reference
        c = np.hstack([c, np.zeros((c.shape[0], c.shape[0] - c.shape[1]))])
    return w, c


def core_guess(h_core, x_large, n_electrons):
    """Core-Hamiltonian starting density (main.cpp:5240-5250)."""
    _, c = _orbitals(h_core, x_large)
    return density_matrix(c, n_electrons)


def commutator_error(fock, density, overlap):
    """scfCommutatorError: F P S - S P F."""
    fps = fock @ density @ overlap
    return fps - fps.T


class Diis:
    """Pulay DIIS with real weights; newest-first history of at most max_vectors pairs; the oldest pair
    is dropped while the scaled bordered system is singular (DIIS.cpp:79-130)."""

    def __init__(self, max_vectors):
        self.max_vectors = max_vectors
        self.errors = deque()
        self.values = deque()

    def extrapolate(self, error, value):
        if self.max_vectors < 2:
            return value
        self.errors.appendleft(np.array(error))
        self.values.appendleft(np.array(value))
        while len(self.errors) > self.max_vectors:
            self.errors.pop()
            self.values.pop()
        e = np.array([x.ravel() for x in self.errors])
        b_full = e @ e.T
        for m in range(len(self.errors), 1, -1):
            scale = np.max(np.diag(b_full)[:m])
            if not scale > 0.0:
                return value
            a = np.zeros((m + 1, m + 1))
            a[:m, :m] = b_full[:m, :m] / scale
            a[:m, m] = a[m, :m] = -1.0
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                lu_ok = np.linalg.cond(a) < 1e14
            except np.linalg.LinAlgError:
                lu_ok = False
            if not lu_ok:
                continue
            w = np.linalg.solve(a, rhs)[:m]
            return sum(wi * vi for wi, vi in zip(w, list(self.values)[:m]))
        return value


@dataclass
class ScfResult:
    converged: bool = False
    iterations: int = 0          # = number of Fock (J/K) builds
    electronic_energy: float = 0.0
    nuclear_repulsion_energy: float = 0.0
    total_energy: float = 0.0
    density_matrix: np.ndarray = None
    fock_matrix: np.ndarray = None
    c_matrix: np.ndarray = None
    orbital_energies: np.ndarray = None
    history: list = field(default_factory=list)


def nonrel_scf(fock_build, h_core, x_large, overlap, initial_density, n_electrons, nuclear_repulsion=0.0,
               diis_size=5, max_iterations=100, energy_tolerance=1e-8, density_tolerance=1e-6, mixing=0.5):
    """nonRelScfImpl (NonRelHartreeFock.cpp:118-198) with `fock_build(P) -> F`; the defaults are the
    reference's input defaults (Input/Input.h:727-730)."""
    res = ScfResult(nuclear_repulsion_energy=nuclear_repulsion)
    p_cur = np.array(initial_density, dtype=np.float64)
    prev = 0.0
    use_diis = diis_size >= 2
    diis = Diis(diis_size if use_diis else 0)
    for it in range(1, max_iterations + 1):
        fock = np.asarray(fock_build(p_cur), dtype=np.float64)
        f_scf = diis.extrapolate(commutator_error(fock, p_cur, overlap), fock) if use_diis else fock
        eps, c = _orbitals(f_scf, x_large)
        p_new = density_matrix(c, n_electrons)
        energy = 0.5 * float(np.sum(p_cur * (h_core + fock)))
        dp = float(np.max(np.abs(p_new - p_cur)))
        de = abs(energy - prev)
        res.iterations = it
        res.electronic_energy = energy
        res.orbital_energies = eps
        res.density_matrix = p_cur
        res.fock_matrix = f_scf
        res.c_matrix = c
        res.history.append((it, energy, dp if it > 1 else float("nan"), de if it > 1 else float("nan")))
        prev = energy
        if it > 1 and (dp < density_tolerance or de < energy_tolerance):
            res.converged = True
            break
        p_cur = p_new if use_diis else mixing * p_new + (1.0 - mixing) * p_cur
    res.total_energy = res.electronic_energy + res.nuclear_repulsion_energy
    return res


@dataclass
class MixedScfResult(ScfResult):
    full_builds: int = 0         # FP64 J/K builds of the whole density
    delta_builds: int = 0        # low-precision incremental builds (J/K of D_new - D_old)
    switch_iteration: int = 0    # first iteration built in FP64 after the incremental phase


def nonrel_scf_incremental(full_build, delta_build, h_core, x_large, overlap, n_electrons, nuclear_repulsion=0.0,
                           initial_orbitals=None, switch_error=1e-5, refresh_factor=1e-1, refresh_every=4, diis_size=5,
                           max_iterations=100, energy_tolerance=1e-8, density_tolerance=1e-6):
    """The NON_REL SCF loop of nonrel_scf (NonRelHartreeFock.cpp:118-198) with incremental Fock builds.

    full_build(C_occ) returns the two-electron part G(D) = J - K/2 of D = 2 C_occ C_occ^T in FP64;
    delta_build(C_new, C_old) returns G(D_new) - G(D_old) in the low precision (FP32 or BF16
    operands: its error is proportional to |D_new - D_old|, not to |D|).

    Every Fock matrix of the incremental phase is G(D_ref) + delta_build(C, C_ref) on top of the last
    full FP64 build G(D_ref), so low-precision errors do not accumulate from one iteration to the
    next. The reference density is refreshed by a full build whenever the largest element of the
    commutator error F D S - S D F has dropped by `refresh_factor` since the last one, or after
    `refresh_every` incremental builds on the same reference; once that error is below
    `switch_error` the loop builds every Fock matrix in full FP64 and converges on those, with the
    reference's OR rule applied only between two FP64 iterations, so the converged energy is the FP64
    fixed point. initial_orbitals: n x n orbitals (occupied first); default the core guess."""
    res = MixedScfResult(nuclear_repulsion_energy=nuclear_repulsion)
    no = n_electrons // 2
    if initial_orbitals is None:
        _, initial_orbitals = _orbitals(h_core, x_large)
    c_cur = np.ascontiguousarray(initial_orbitals[:, :no])
    c_ref, g_ref = c_cur, np.asarray(full_build(c_cur), dtype=np.float64)
    g, g_kind, err_ref, since_ref = g_ref, "f64", None, 0
    res.full_builds = 1
    mode_full, full_since_switch = False, 0
    diis = Diis(diis_size if diis_size >= 2 else 0)
    prev = 0.0
    for it in range(1, max_iterations + 1):
        p_cur = 2.0 * c_cur @ c_cur.T
        fock = h_core + g
        err = commutator_error(fock, p_cur, overlap)
        emax = float(np.max(np.abs(err)))
        if err_ref is None:
            err_ref = emax
        f_scf = diis.extrapolate(err, fock) if diis_size >= 2 else fock
        eps, c = _orbitals(f_scf, x_large)
        c_new = np.ascontiguousarray(c[:, :no])
        p_new = 2.0 * c_new @ c_new.T
        energy = 0.5 * float(np.sum(p_cur * (h_core + fock)))
        dp = float(np.max(np.abs(p_new - p_cur)))
        de = abs(energy - prev)
        res.iterations, res.electronic_energy, res.orbital_energies = it, energy, eps
        res.density_matrix, res.fock_matrix, res.c_matrix = p_cur, f_scf, c
        res.history.append((it, energy, dp if it > 1 else float("nan"), de if it > 1 else float("nan"), g_kind))
        if mode_full and full_since_switch >= 2 and (dp < density_tolerance or de < energy_tolerance):
            res.converged = True
            break
        prev = energy
        if it == max_iterations:
            break
        if not mode_full and emax < switch_error:
            mode_full = True
            res.switch_iteration = it + 1
        if mode_full or emax < refresh_factor * err_ref or since_ref >= refresh_every:
            g = g_ref = np.asarray(full_build(c_new), dtype=np.float64)
            c_ref, err_ref, g_kind, since_ref = c_new, None, "f64", 0
            res.full_builds += 1
            full_since_switch += int(mode_full)
        else:
            g = g_ref + np.asarray(delta_build(c_new, c_ref), dtype=np.float64)
            res.delta_builds += 1
            g_kind, since_ref = "delta", since_ref + 1
        c_cur = c_new
    res.total_energy = res.electronic_energy + res.nuclear_repulsion_energy
    return res


def gpu_builds(ctx, slot, precision):
    """(full_build, delta_build) for nonrel_scf_incremental on the GPU: rr_ri_jk_factor with
    L = sqrt(2) C_occ in FP64 and rr_ri_jk_delta in `precision`."""
    from . import FP64

    def full(c_occ):
        j, k = ctx.ri_jk_factor(slot, None, np.sqrt(2.0) * c_occ, precision=FP64)
        return j - 0.5 * k

    def delta(c_new, c_old):
        dj, dk = ctx.ri_jk_delta(slot, c_new, c_old, precision=precision)
        return dj - 0.5 * dk

    return full, delta
