"""4c Dirac-Hartree-Fock in the restricted-kinetic-balance (RKB) basis: the FP64 inputs of the RKB RI
Fock build (rkbFockMatrix with a RiDirectEriSource) from a Gaussian basis, and the DHF SCF loop around
the GPU build.

The inputs are prepared as the reference's C4_SPINOR construction prepares them
(src/main.cpp:4493-4600 and the RI source of :5005-5030):
- small-component functions: for each (normalized, Cartesian) large function its own d/dx, d/dy,
  d/dz pieces in the order lower_x, raise_x, lower_y, raise_y, lower_z, raise_z, a lowering piece only
  where the Cartesian power along the axis is positive (buildRkbSmallBasis,
  src/RKB/RkbDerivativeTerms.cpp:100-132; d/dx of x^j exp(-a r^2) = j x^(j-1) - 2 a x^(j+1), the same
  exponents; not renormalized);
- the unrestricted-kinetic-balance (UKB) Dirac Hamiltonian over (L alpha, L beta, S alpha, S beta):
  c sigma.p between the large and small functions from <L|grad|S> (diracKineticMatrix,
  src/Utils/DiracKinetic.cpp:8-49; p = -i grad), -2 c^2 S_SS on the small blocks
  (diracRestEnergyMatrix, :51-70), the point-nucleus attraction on every block (vextMatrix,
  src/Vext/Vext.cpp:8-40), c = 137.036 (src/Utils/PhysicalConstants.h:10);
- the RKB coefficients C (2 nL x 2 nS: -i D_z, -i D_x + D_y, -i D_x - D_y, +i D_z per spin pair,
  rkbCoefficients, src/RKB/RkbTransformation.cpp:19-61) and the embedding V (identity on the large
  spinors, V[2nL + t][2nL + j] = C[j][t], rkbEmbeddingMatrix, src/RKB/RkbHamiltonian.cpp:8-27);
  h_RKB = V^H h_UKB V; the RKB small overlap conj(C) blockdiag(S_SS, S_SS) C^T
  (src/RKB/RkbOverlap.cpp:10-48); S_full, X_full = blockdiag of the large Loewdin orthogonalizer (twice)
  and the inverse square root of the RKB small overlap (src/RKB/RkbOrthogonalization.cpp);
- the RI auxiliary basis: PAUTO over the union of the large and small functions' unique
  (centre, l, exponents) shells, metric canonically orthogonalized at 1e-10, B_LL and B_SS over the
  Cartesian large and small functions sharing that aux index (main.cpp:5008-5027).
The large component is kept Cartesian, which equals the reference's spherical large basis for s and p
shells (the only shells accepted here).

The SCF is dhfScfImpl (src/C4_DHF/C4_DHF.cpp:37-142): F = rkbFockMatrix(h, eri, P), DIIS on the
commutator error, diagonalisation of X^H F X, the density of the n_electrons lowest positive-energy
spinors (rkbDensityMatrix, RkbDensityMatrix.cpp:15-40: columns n/2 .. n/2 + n_e - 1), the energy
1/2 Re tr(P (h + F)) with the density that built F, and the OR convergence rule. The starting
density is the reference's: the same occupation of the eigenvectors of X^H h_RKB X
(main.cpp:6117-6119, x2cDecoupling).

Integrals are evaluated over a pool of single-primitive Cartesian shells (every primitive of every
large and small function) and contracted to the functions, so the small functions, which are single
Cartesian components with their own derivative weights, need no shell structure of their own.
"""
import numpy as np

from . import integrals as ig

SPEED_OF_LIGHT = 137.036


def _inv_sqrt_h(m):
    w, v = np.linalg.eigh(m)
    return (v / np.sqrt(w)) @ v.conj().T


class DiracMolecule:
    """A molecule in a Gaussian basis prepared for the 4c RKB RI SCF.

    atoms: [(symbol, xyz in bohr)]; basis: integrals.parse_gbs result (s and p shells)."""

    def __init__(self, atoms, basis, n_electrons=None, speed_of_light=SPEED_OF_LIGHT, aux_kind="PAUTO",
                 metric_threshold=1e-10):
        self.symbols = [a[0] for a in atoms]
        self.charges = np.array([float(ig.ELEMENTS[a[0]]) for a in atoms])
        self.positions = np.array([a[1] for a in atoms], dtype=np.float64).reshape(-1, 3)
        self.shells = ig.molecule_shells(list(zip(self.symbols, self.positions)), basis)
        if any(s.l > 1 for s in self.shells):
            raise ValueError("DiracMolecule: large shells above p need the spherical large-component reduction")
        self.c = float(speed_of_light)
        self.aux_kind = aux_kind
        self.metric_threshold = metric_threshold
        self.n_electrons = int(round(self.charges.sum())) if n_electrons is None else int(n_electrons)
        self._ints = None

    # ---- functions -----------------------------------------------------------------------------
    def _functions(self):
        """Large functions (centre, comp, exps, literal coefficients), small functions likewise, and
        each large function's term indices {lower_x: j, raise_x: j, ...} into the small list."""
        s_raw = ig.one_electron(self.shells, [], np.zeros((0, 3)))[0]
        nrm = ig.normalization(np.diag(s_raw))
        large, k = [], 0
        for s in self.shells:
            for comp in ig.cart_components(s.l):
                d = nrm[k] * s.coefs * ig._prim_norm(s.exps, s.l) * ig._angular_factor(comp)
                large.append((s.centre, comp, s.exps, d, s.element))
                k += 1
        small, terms = [], []
        for ctr, comp, exps, d, el in large:
            t = {}
            for axis, name in ((0, "x"), (1, "y"), (2, "z")):
                if comp[axis] > 0:
                    lc = list(comp)
                    lc[axis] -= 1
                    t["lower_" + name] = len(small)
                    small.append((ctr, tuple(lc), exps, comp[axis] * d, el))
                rc = list(comp)
                rc[axis] += 1
                t["raise_" + name] = len(small)
                small.append((ctr, tuple(rc), exps, -2.0 * exps * d, el))
            terms.append(t)
        return large, small, terms

    @staticmethod
    def _pool(funcs_sets):
        """Single-primitive shells covering every primitive of the given functions, and per set the
        matrix T (n_functions x n_pool) with function = sum_q T[., q] pool_q (pool functions in the
        engine's convention: a unit-coefficient primitive shell's component carries its angular factor)."""
        keys = {}
        for funcs in funcs_sets:
            for ctr, comp, exps, _, el in funcs:
                l = sum(comp)
                for a in exps:
                    keys.setdefault((tuple(float(v) for v in ctr), l, float(a)), el)
        shells, offset, n = [], {}, 0
        for key, el in keys.items():
            ctr, l, a = key
            shells.append(ig.Shell(np.array(ctr), l, [a], [1.0 / ig._prim_norm(a, l)], el))
            offset[key] = n
            n += shells[-1].size
        mats = []
        for funcs in funcs_sets:
            t = np.zeros((len(funcs), n))
            for i, (ctr, comp, exps, d, _) in enumerate(funcs):
                l = sum(comp)
                col = ig.cart_components(l).index(tuple(comp))
                for a, dk in zip(exps, d):
                    t[i, offset[(tuple(float(v) for v in ctr), l, float(a))] + col] += dk / ig._angular_factor(comp)
            mats.append(t)
        return shells, mats

    def _aux_shells(self):
        """PAUTO (or AUTO) over the unique (centre, l, exponents) shells of the large and small
        functions (buildAutoAuxiliaryBasis deduplicates on exactly that key)."""
        pseudo = []
        for s in self.shells:
            pseudo.append(s)
            pseudo.append(ig.Shell(s.centre, s.l + 1, s.exps, s.coefs, s.element))
            if s.l >= 1:
                pseudo.append(ig.Shell(s.centre, s.l - 1, s.exps, s.coefs, s.element))
        return ig.auto_aux_shells(pseudo, self.aux_kind)

    # ---- inputs --------------------------------------------------------------------------------
    def integrals(self):
        """dict: h_rkb, s_full, x_full, v_total (complex), eri3_ll, eri3_ss ((n_aux, n, n) real),
        nl, ns, n_rkb, e_nuc, n_aux_raw, n_aux_dropped, s_large, x_large."""
        if self._ints is not None:
            return dict(self._ints)
        c = self.c
        large, small, terms = self._functions()
        nl, ns = len(large), len(small)
        pool, (tl, ts) = self._pool([large, small])
        s_p, _, v_p = ig.one_electron(pool, self.charges, self.positions)
        d_p = ig.nabla(pool)
        s_l, v_l = tl @ s_p @ tl.T, tl @ v_p @ tl.T
        s_s, v_s = ts @ s_p @ ts.T, ts @ v_p @ ts.T
        dx, dy, dz = (tl @ d_p[k] @ ts.T for k in range(3))
        # UKB Dirac Hamiltonian, blocks (La, Lb, Sa, Sb)
        nu = 2 * nl + 2 * ns
        la, lb, sa, sb = 0, nl, 2 * nl, 2 * nl + ns
        h = np.zeros((nu, nu), dtype=np.complex128)
        f = -1j * c
        for r0, c0, blk in ((la, sa, f * dz), (la, sb, f * (dx - 1j * dy)), (lb, sa, f * (dx + 1j * dy)), (lb, sb, -f * dz)):
            h[r0:r0 + nl, c0:c0 + ns] = blk
            h[c0:c0 + ns, r0:r0 + nl] = blk.conj().T
        for o in (la, lb):
            h[o:o + nl, o:o + nl] += v_l
        for o in (sa, sb):
            h[o:o + ns, o:o + ns] += v_s - 2.0 * c * c * s_s
        # RKB coefficients and embedding
        cr = np.zeros((2 * nl, 2 * ns), dtype=np.complex128)
        for p, t in enumerate(terms):
            ra, rb = p, nl + p

            def add(row, off, names, w):
                for nm in names:
                    if nm in t:
                        cr[row, off + t[nm]] += w
            add(ra, 0, ("lower_z", "raise_z"), -1j)
            add(ra, ns, ("lower_x", "raise_x"), -1j)
            add(ra, ns, ("lower_y", "raise_y"), 1.0)
            add(rb, 0, ("lower_x", "raise_x"), -1j)
            add(rb, 0, ("lower_y", "raise_y"), -1.0)
            add(rb, ns, ("lower_z", "raise_z"), 1j)
        n_rkb = 4 * nl
        v = np.zeros((nu, n_rkb), dtype=np.complex128)
        v[:2 * nl, :2 * nl] = np.eye(2 * nl)
        v[2 * nl:, 2 * nl:] = cr.T
        h_rkb = v.conj().T @ h @ v
        s_small_ukb = np.zeros((2 * ns, 2 * ns))
        s_small_ukb[:ns, :ns] = s_s
        s_small_ukb[ns:, ns:] = s_s
        s_small = cr.conj() @ s_small_ukb @ cr.T
        x_l = _inv_sqrt_h(s_l)
        x_small = _inv_sqrt_h(s_small)
        s_full = np.zeros((n_rkb, n_rkb), dtype=np.complex128)
        x_full = np.zeros((n_rkb, n_rkb), dtype=np.complex128)
        for o in (0, nl):
            s_full[o:o + nl, o:o + nl] = s_l
            x_full[o:o + nl, o:o + nl] = x_l
        s_full[2 * nl:, 2 * nl:] = s_small
        x_full[2 * nl:, 2 * nl:] = x_small
        # RI tensors over one shared aux index
        aux = self._aux_shells()
        an = ig.normalization(ig.self_overlap_prim(aux))
        raw = ig.three_center_prim(pool, aux) * an[:, None, None]
        metric = ig.two_center_prim(aux) * np.outer(an, an)
        w, u = np.linalg.eigh(metric)
        keep = w > self.metric_threshold
        xm = u[:, keep] / np.sqrt(w[keep])

        def contract(t):
            r = np.einsum("qab,ia,jb->qij", raw, t, t, optimize=True)
            b = np.einsum("qij,qp->pij", r, xm, optimize=True)
            return np.ascontiguousarray(0.5 * (b + np.transpose(b, (0, 2, 1))))

        self._ints = {"h_rkb": h_rkb, "s_full": s_full, "x_full": x_full, "v_total": v, "eri3_ll": contract(tl),
                      "eri3_ss": contract(ts), "nl": nl, "ns": ns, "n_rkb": n_rkb,
                      "e_nuc": ig.nuclear_repulsion(self.charges, self.positions), "n_aux_raw": metric.shape[0],
                      "n_aux_dropped": int(np.sum(~keep)), "s_large": s_l, "x_large": x_l, "h_ukb": h,
                      "rkb_coefficients": cr}
        return dict(self._ints)


# ---- SCF ---------------------------------------------------------------------------------------
def rkb_density(c, n_electrons):
    """rkbDensityMatrix: sum over the spinors n/2 .. n/2 + n_e - 1 of c c^H (the n/2 lowest are the
    negative-energy branch)."""
    n = c.shape[0]
    if c.shape[1] != n or n % 2:
        raise ValueError("rkbDensityMatrix: C_DHF must be square with an even dimension")
    if n_electrons <= 0 or n // 2 + n_electrons > n:
        raise ValueError("rkbDensityMatrix: n_electrons must fit within the positive-energy half")
    co = c[:, n // 2:n // 2 + n_electrons]
    return co @ co.conj().T


def _spinors(f, x):
    w, u = np.linalg.eigh(x.conj().T @ f @ x)
    return w, x @ u


def core_guess(h_rkb, x_full, n_electrons):
    """Starting density of the 4c SCF: the eigenvectors of X^H h_RKB X (main.cpp:6117-6119)."""
    return rkb_density(_spinors(h_rkb, x_full)[1], n_electrons)


class DhfResult:
    def __init__(self):
        self.iterations = 0
        self.converged = False
        self.electronic_energy = 0.0
        self.total_energy = 0.0
        self.nuclear_repulsion_energy = 0.0
        self.density_matrix = None
        self.fock_matrix = None
        self.orbital_energies = None
        self.history = []


def dhf_scf(fock_build, h_rkb, x_full, s_full, initial_density, n_electrons, nuclear_repulsion=0.0, diis=None,
            max_iterations=100, energy_tolerance=1e-8, density_tolerance=1e-6):
    """dhfScfImpl (C4_DHF.cpp:37-142) with `fock_build(P) -> F` and `diis(error, F) -> F` (None: plain
    iteration). The error is the commutator F P S - S P F (scfCommutatorError, DIIS.cpp:51-58)."""
    res = DhfResult()
    res.nuclear_repulsion_energy = nuclear_repulsion
    p_cur = np.array(initial_density, dtype=np.complex128)
    prev = 0.0
    for it in range(1, max_iterations + 1):
        fock = np.asarray(fock_build(p_cur), dtype=np.complex128)
        if diis is not None:
            fps = fock @ p_cur @ s_full
            err = fps - s_full @ p_cur @ fock
            f_scf = diis(err, fock)
        else:
            f_scf = fock
        eps, c = _spinors(f_scf, x_full)
        p_new = rkb_density(c, n_electrons)
        energy = 0.5 * float(np.real(np.sum(p_cur * (h_rkb + fock).T)))
        dp = float(np.max(np.abs(p_new - p_cur)))
        de = abs(energy - prev)
        res.iterations, res.electronic_energy, res.orbital_energies = it, energy, eps
        res.density_matrix, res.fock_matrix = p_cur, f_scf
        res.history.append((it, energy, dp if it > 1 else float("nan"), de if it > 1 else float("nan")))
        prev = energy
        if it > 1 and (dp < density_tolerance or de < energy_tolerance):
            res.converged = True
            break
        p_cur = p_new
    res.total_energy = res.electronic_energy + res.nuclear_repulsion_energy
    return res


def geometry_inputs(args):
    """integrals() of DiracMolecule(atoms, basis, **options) for args = (atoms, basis, options): a
    top-level function, so that the host integrals of many geometries can be evaluated in worker
    processes."""
    atoms, basis, options = args
    try:  # one BLAS thread per worker: the workers themselves are the parallelism
        from threadpoolctl import threadpool_limits
    except ImportError:
        return DiracMolecule(atoms, basis, **options).integrals()
    with threadpool_limits(1):
        return DiracMolecule(atoms, basis, **options).integrals()


def run_dhf(ctx, mol, initial_density=None, slot_ll=0, slot_ss=1, diis_size=8, **scf_options):
    """4c RKB RI SCF of a DiracMolecule with the GPU Fock build (rkbFockMatrix, RI source:
    recovery of the occupied spinors in the RKB basis, the four-flavour build on the UKB tensors,
    V^H G V). Returns (DhfResult, ints)."""
    ints = mol.integrals()
    return run_dhf_inputs(ctx, ints, mol.n_electrons, initial_density, slot_ll, slot_ss, diis_size, **scf_options), ints


def run_dhf_inputs(ctx, ints, n_electrons, initial_density=None, slot_ll=0, slot_ss=1, diis_size=8, **scf_options):
    """run_dhf on prepared inputs (DiracMolecule.integrals()); returns the DhfResult."""
    ctx.upload(slot_ll, ints["eri3_ll"])
    ctx.upload(slot_ss, ints["eri3_ss"])
    h, s, x, v = ints["h_rkb"], ints["s_full"], ints["x_full"], ints["v_total"]
    p0 = core_guess(h, x, n_electrons) if initial_density is None else initial_density
    handle = ctx.diis_create(diis_size) if diis_size >= 2 else None
    try:
        diis = (lambda e, f: ctx.diis_extrapolate(handle, e, f)[0]) if handle is not None else None
        res = dhf_scf(lambda p: ctx.rkb_fock_ri(slot_ll, slot_ss, ints["nl"], ints["ns"], h, v, s, x, p), h, x, s, p0,
                      n_electrons, ints["e_nuc"], diis=diis, **scf_options)
    finally:
        if handle is not None:
            ctx.diis_destroy(handle)
    return res
