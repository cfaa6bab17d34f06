"""Seeded synthetic inputs for the RI J/K workloads (host side, numpy).

Shapes and distributions follow the benchmark configurations of the RI Fock-build path:
  * kernel-level water/cc-pVDZ: B[300][24][24] symmetric, N(0, 0.1) * exp(-|mu-nu|/8), 5 occupied
    orbitals from the QR factor of an N(0,1) 24 x 24 matrix;
  * large real system: B[4000][1000][1000] symmetric, N(0,1) * exp(-|mu-nu|/50), 200 occupied;
  * X2C 2-spinor: 400 spatial functions, 2400 aux, 160 occupied Kramers-paired spinors;
  * 4c RKB: nL = 300, nS = 700, shared aux index of 4200 (the LL block non-zero on the first 1800).
The default seed of configuration i is 20261203 + i. Large tensors are generated on the device
instead (Context.generate); this module is for the sizes a CPU check can handle.
"""
import numpy as np

BASE_SEED = 20261203


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def decay_matrix(n, length):
    idx = np.arange(n)
    return np.exp(-np.abs(idx[:, None] - idx[None, :]) / float(length))


def ri_tensor(naux, n, sigma, decay, seed, symmetric=True):
    """(naux, n, n) real tensor, sigma * N(0,1) * exp(-|mu-nu|/decay), symmetric in mu nu."""
    g = rng(seed)
    z = g.standard_normal((naux, n, n))
    if symmetric:
        iu = np.triu_indices(n, 1)
        z[:, iu[1], iu[0]] = z[:, iu[0], iu[1]]
    return sigma * z * decay_matrix(n, decay)[None, :, :]


def hermitian_ri_tensor(naux, n, sigma, decay, seed):
    """(naux, n, n) complex Hermitian tensor with re/im parts sigma * N(0,1) * exp(-|mu-nu|/decay)."""
    g = rng(seed)
    re = g.standard_normal((naux, n, n))
    im = g.standard_normal((naux, n, n))
    z = re + 1j * im
    z = 0.5 * (z + np.conj(np.transpose(z, (0, 2, 1))))
    return sigma * z * decay_matrix(n, decay)[None, :, :]


def orthonormal_occ(n, nocc, seed):
    """First nocc columns of the Q factor of an N(0,1) n x n matrix."""
    q, _ = np.linalg.qr(rng(seed).standard_normal((n, n)))
    return np.ascontiguousarray(q[:, :nocc])


def rotate_occ(c_occ, n_total_basis_q, angle, seed):
    """Rotate an orthonormal occupied set by exp(kappa) with kappa an occ-virt generator of norm angle."""
    n, no = c_occ.shape
    g = rng(seed)
    q = n_total_basis_q
    virt = q[:, no:]
    kappa = g.standard_normal((virt.shape[1], no))
    kappa *= angle / np.linalg.norm(kappa)
    # exact rotation within span[occ, virt] via the SVD of kappa
    u, s, vt = np.linalg.svd(kappa, full_matrices=False)
    c = c_occ @ vt.T
    v = virt @ u
    rot = c * np.cos(s)[None, :] + v * np.sin(s)[None, :]
    return np.ascontiguousarray(rot @ vt)


def density(c_occ, factor=2.0):
    return factor * c_occ @ c_occ.T


def symmetric(n, seed, scale=1.0):
    a = rng(seed).standard_normal((n, n)) * scale
    return 0.5 * (a + a.T)


def spd_overlap(n, seed, offdiag=0.05):
    """A well-conditioned SPD 'overlap' with unit diagonal."""
    a = rng(seed).standard_normal((n, n)) * offdiag
    s = np.eye(n) + 0.5 * (a + a.T)
    w = np.linalg.eigvalsh(s)
    if w.min() < 0.2:
        s += (0.2 - w.min()) * np.eye(n)
    d = 1.0 / np.sqrt(np.diag(s))
    return s * d[:, None] * d[None, :]


def inverse_sqrt(s):
    w, v = np.linalg.eigh(s)
    return (v / np.sqrt(w)) @ v.T


def s_orthonormal_occ(s, nocc, seed):
    """C with C^T S C = I (first nocc columns of X Q, X = S^-1/2)."""
    x = inverse_sqrt(s)
    q = orthonormal_occ(s.shape[0], nocc, seed)
    return x @ q


def hermitian(n, seed, scale=1.0):
    g = rng(seed)
    a = (g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))) * scale
    return 0.5 * (a + a.conj().T)


def unitary(n, seed):
    g = rng(seed)
    q, r = np.linalg.qr(g.standard_normal((n, n)) + 1j * g.standard_normal((n, n)))
    return q * (np.diag(r) / np.abs(np.diag(r)))[None, :]


def kramers_pair_occ(nl, npairs, seed):
    """Kramers-paired spinor coefficients in [alpha; beta] blocks: (phi_a, phi_b) and its partner
    (-conj(phi_b), conj(phi_a)); columns orthonormal. Returns 2nl x 2npairs."""
    g = rng(seed)
    a = g.standard_normal((2 * nl, npairs)) + 1j * g.standard_normal((2 * nl, npairs))
    cols = []
    for k in range(npairs):
        v = a[:, k].copy()
        for w in cols:
            v -= w * (w.conj() @ v)
        pa, pb = v[:nl], v[nl:]
        partner = np.concatenate([-np.conj(pb), np.conj(pa)])
        for w in cols:
            partner -= w * (w.conj() @ partner)
        partner -= v * ((v.conj() @ partner) / (v.conj() @ v))
        v /= np.linalg.norm(v)
        partner /= np.linalg.norm(partner)
        cols.extend([v, partner])
    return np.stack(cols, axis=1)


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def device_tensor_rows(n, sigma, decay, seed, p0, p1):
    """Rows [p0, p1) of the tensor Context.generate(slot, n_aux, n, sigma, decay, seed) builds on the
    device, (p1 - p0, n, n) float64: value(P, mu, nu) = sigma z exp(-|mu - nu| / decay) with z the
    Box-Muller normal of two splitmix64 draws keyed by (seed, P, min(mu, nu), max(mu, nu)). The rows
    depend on P alone, so a tensor generated with fewer aux rows is the head of a larger one."""
    iu = np.triu_indices(n)
    a, b = iu[0].astype(np.uint64), iu[1].astype(np.uint64)
    dec = sigma * np.exp(-(iu[1] - iu[0]).astype(np.float64) / float(decay))
    out = np.empty((p1 - p0, n, n))
    s = np.uint64(seed)
    with np.errstate(over="ignore"):
        for P in range(p0, p1):
            key = _splitmix64(s ^ _splitmix64((np.uint64(P) * np.uint64(n) + a) * np.uint64(n) + b))
            k2 = _splitmix64(key)
            u1 = ((key >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740993.0)
            u2 = (k2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
            v = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2) * dec
            blk = out[P - p0]
            blk[iu] = v
            blk[iu[1], iu[0]] = v
    return out


def kramers_pair_ukb(nl, ns, npairs, seed):
    """Kramers-paired occupied spinors in the four UKB flavor blocks [La; Lb; Sa; Sb] (sizes nl, nl, ns,
    ns): each vector v comes with its time-reversed partner (-conj(v_Lb), conj(v_La), -conj(v_Sb),
    conj(v_Sa)); columns orthonormal. Returns (2nl + 2ns) x 2npairs."""
    g = rng(seed)
    n = 2 * nl + 2 * ns

    def partner(v):
        la, lb, sa, sb = v[:nl], v[nl:2 * nl], v[2 * nl:2 * nl + ns], v[2 * nl + ns:]
        return np.concatenate([-np.conj(lb), np.conj(la), -np.conj(sb), np.conj(sa)])

    cols = []
    a = g.standard_normal((n, npairs)) + 1j * g.standard_normal((n, npairs))
    for k in range(npairs):
        v = a[:, k].copy()
        for _ in range(2):
            for w in cols:
                v -= w * (w.conj() @ v)
        v /= np.linalg.norm(v)
        t = partner(v)  # orthogonal to v and to every earlier pair (time reversal is antiunitary)
        t /= np.linalg.norm(t)
        cols.extend([v, t])
    return np.stack(cols, axis=1)


def time_reversal_ukb(nl, ns):
    """Unitary part U of the UKB time-reversal operator Theta = U K: (La, Lb, Sa, Sb) -> (-Lb, La, -Sb, Sa).
    A time-reversal symmetric operator satisfies G = U conj(G) U^H."""
    n = 2 * nl + 2 * ns
    u = np.zeros((n, n))
    for off, m in ((0, nl), (2 * nl, ns)):
        u[off:off + m, off + m:off + 2 * m] = -np.eye(m)
        u[off + m:off + 2 * m, off:off + m] = np.eye(m)
    return u
