"""GPU parity of the functional stage (riHxFockGemm, the sigma toolkit, pair/diag contractions,
accumU and the four sigma-vector drivers) against the oracle, through the C ABI.

Tolerance: FP64 throughout; 1e-11 relative to the largest entry (different summation order).
"""
import numpy as np
import pytest

import rerdmft_amd as rr

from oracle import oracle  # noqa: E402

import _dense as dn  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-11


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def case(seed, n, naux, cplx, hermitian=True):
    g = np.random.default_rng(seed)
    b = g.standard_normal((naux, n, n)) * np.exp(-np.abs(np.subtract.outer(np.arange(n), np.arange(n))) / 6.0)
    if cplx:
        b = b + 1j * g.standard_normal((naux, n, n))
    if hermitian:
        b = 0.5 * (b + np.conj(np.transpose(b, (0, 2, 1))))
    h = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0.0)
    h = 0.5 * (h + np.conj(h.T))
    occ = g.uniform(0.05, 0.95, n)
    h2, x2 = g.standard_normal((n, n)), g.standard_normal((n, n))  # unequal and asymmetric on purpose
    return g, b, h, occ, h2, x2


def pairing(n):
    """Adjacent pairs (p <-> p xor 1); an unpaired last orbital is its own partner."""
    return np.array([i ^ 1 if (i ^ 1) < n else i for i in range(n)])


@pytest.fixture(params=[None, "300"], ids=["one_chunk", "many_chunks"])
def chunking(request, monkeypatch):
    if request.param:
        monkeypatch.setenv("RR_SIGMA_CHUNK_ELEMS", request.param)
    return request.param


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("with_l", [False, True])
@pytest.mark.parametrize("shape", [(12, 30), (33, 70)])
def test_hx_fock_gemm(ctx, cplx, with_l, shape, chunking):
    n, naux = shape
    g, b, h, occ, h2, x2 = case(100 + n + cplx, n, naux, cplx, hermitian=False)
    pr = l1 = l2 = None
    if with_l:
        pr, l1, l2 = pairing(n), g.standard_normal((n, n)), g.standard_normal((n, n))
    f = ctx.ri_hx_fock_gemm(b, h, occ, h2, x2, pr, l1, l2)
    fr = oracle.ri_hx_fock_gemm(b.reshape(naux, -1), h, occ, h2, x2, pr, l1, l2)
    assert rel(f, fr) < TOL


def test_hx_fock_gemm_matches_dense_definition(ctx):
    g, b, h, occ, h2, x2 = case(7, 6, 9, True)
    pr = pairing(6)
    l1, l2 = g.standard_normal((6, 6)), g.standard_normal((6, 6))
    f = ctx.ri_hx_fock_gemm(b, h, occ, h2, x2, pr, l1, l2)
    assert rel(f, dn.fock(h, b, occ, h2, x2, pair=pr, l1=l1, l2=l2)) < TOL


def test_hx_fock_gemm_shape_error(ctx):
    g, b, h, occ, h2, x2 = case(8, 6, 9, False)
    with pytest.raises(rr.GpuError, match="pair_of"):
        ctx.ri_hx_fock_gemm(b, h, occ, h2, x2, pairing(6), None, None)


@pytest.mark.parametrize("cplx", [False, True])
def test_toolkit(ctx, cplx):
    g, b, h, occ, h2, x2 = case(9 + cplx, 21, 40, cplx)
    k = g.standard_normal((21, 21))
    k = k - k.T
    got = ctx.ri_sigma_toolkit(b.reshape(40, -1), k)
    want = oracle.ri_sigma_toolkit(b.reshape(40, -1), k)
    for a, w in zip(got, want):
        assert rel(a.reshape(w.shape), w) < TOL


@pytest.mark.parametrize("cplx", [False, True])
def test_pair_diag_accum_primitives(ctx, cplx, chunking):
    n, naux = 17, 29
    g, b, h, occ, h2, x2 = case(11 + cplx, n, naux, cplx, hermitian=False)
    bf = b.reshape(naux, -1)
    idx_f = g.permutation(n * n).reshape(n, n).astype(np.uint32)
    idx_g = np.add.outer(np.arange(n) * n, np.arange(n)).T.astype(np.uint32).copy()
    row = g.standard_normal((naux, n)) + (1j * g.standard_normal((naux, n)) if cplx else 0.0)
    f = dict(s=bf, idx=idx_f, w=g.standard_normal((n, n)), row=row, post=g.standard_normal((n, n)))
    gg = dict(s=bf[::-1].copy(), idx=idx_g, w=g.standard_normal((n, n)), post=g.standard_normal((n, n)))
    t0 = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0.0)
    got = ctx.ri_pair_contract(t0.copy(), -1.5, n, naux, f, gg)
    want = oracle.ri_pair_contract(t0.copy(), -1.5, n, naux, f, gg)
    assert rel(got, want) < TOL
    # operand without a source tensor (weights and row factors only)
    f2 = dict(w=g.standard_normal((n, n)), row=row)
    got = ctx.ri_pair_contract(t0.copy(), 1.0, n, naux, f2, dict(s=bf, idx=idx_f))
    want = oracle.ri_pair_contract(t0.copy(), 1.0, n, naux, f2, dict(s=bf, idx=idx_f))
    assert rel(got, want) < TOL
    ua, ub = row, row[::-1].copy()
    assert rel(ctx.ri_diag_contract(t0.copy(), b, ua, ub), oracle.ri_diag_contract(t0.copy(), b, ua, ub)) < TOL
    assert rel(ctx.ri_diag_contract(t0.copy(), b, None, ub), oracle.ri_diag_contract(t0.copy(), b, None, ub)) < TOL
    wm = g.standard_normal((n, n))
    assert rel(ctx.ri_accum_u(ua.copy(), 0.25, ub, wm), oracle.ri_accum_u(ua.copy(), 0.25, ub, wm)) < TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("kind", ["jk", "hx", "hx_l"])
def test_sigma_drivers(ctx, cplx, kind, chunking):
    n, naux = 14, 31
    g, b, h, occ, h2, x2 = case(200 + cplx + 3 * len(kind), n, naux, cplx)
    pairs = dn.lower_pairs(n)
    v = g.standard_normal(len(pairs) * (2 if cplx else 1))
    if kind == "jk":
        fn = ctx.jk_only_sigma_joint_hessian_vector_ri if cplx else ctx.jk_only_sigma_hessian_vector_ri
        w = fn(h, b, occ, h2, x2, pairs, v)
        wr = oracle.sigma_hessian_vector_ri(h, b, occ, h2, x2, pairs, v)
    else:
        pr = l1 = l2 = None
        if kind == "hx_l":
            pr, l1, l2 = pairing(n), g.standard_normal((n, n)), g.standard_normal((n, n))
        f = oracle.ri_hx_fock_gemm(b.reshape(naux, -1), h, occ, h2, x2, pr, l1, l2)
        fn = ctx.hartree_exchange_sigma_joint_hessian_vector_ri if cplx else ctx.hartree_exchange_sigma_hessian_vector_ri
        w = fn(h, b, occ, h2, x2, f, pairs, v, pr, l1, l2)
        wr = oracle.sigma_hessian_vector_ri(h, b, occ, h2, x2, pairs, v, fock=f, pair_of=pr, l1=l1, l2=l2, hx=True)
    assert rel(w, wr) < 1e-10


def test_jk_sigma_is_the_energy_hessian_on_gpu(ctx):
    """The GPU product against the finite-difference Hessian of the energy (no oracle involved)."""
    g, b, h, occ, h2, x2 = case(300, 5, 6, False)
    h2, x2 = 0.5 * (h2 + h2.T), 0.5 * (x2 + x2.T)
    pairs = dn.lower_pairs(5)
    m = len(pairs)
    H = np.zeros((m, m))
    for j in range(m):
        e = np.zeros(m)
        e[j] = 1.0
        kap = dn.kappa_real(5, pairs, e)
        H[:, j] = dn.d_dt(lambda t: dn.gradient_vector(dn.fock(*dn.rotate(h, b, t * kap), occ, h2, x2), pairs))
    hs = 0.5 * (H + H.T)
    v = g.standard_normal(m)
    w = ctx.jk_only_sigma_hessian_vector_ri(h, b, occ, h2, x2, pairs, v)
    assert np.max(np.abs(w - hs @ v)) < 1e-6 * max(1.0, np.max(np.abs(hs @ v)))


def test_sigma_rejects_bad_pairs(ctx):
    g, b, h, occ, h2, x2 = case(301, 5, 6, False)
    with pytest.raises(rr.GpuError, match="pair index out of range"):
        ctx.jk_only_sigma_hessian_vector_ri(h, b, occ, h2, x2, [(7, 1)], np.ones(1))


def test_fused_functional_fock_rkb(ctx):
    """riJkOnlyFockMatrix / riPnofFockMatrix body on the device vs the oracle's pieces composed."""
    from rerdmft_amd import synthetic as syn
    g = np.random.default_rng(400)
    nl, ns, naux, n_rkb = 9, 7, 40, 20
    nu = 2 * nl + 2 * ns
    bl = syn.ri_tensor(naux, nl, 0.5, 5.0, 401)
    bs = syn.ri_tensor(naux, ns, 0.5, 5.0, 402)
    ctx.upload(10, bl)
    ctx.upload(11, bs)
    v = g.standard_normal((nu, n_rkb)) + 1j * g.standard_normal((nu, n_rkb))
    c = g.standard_normal((n_rkb, 12)) + 1j * g.standard_normal((n_rkb, 12))
    hr = g.standard_normal((n_rkb, n_rkb)) + 1j * g.standard_normal((n_rkb, n_rkb))
    hr = 0.5 * (hr + hr.conj().T)
    occ = g.uniform(0, 1, 12)
    h2, x2 = g.standard_normal((12, 12)), g.standard_normal((12, 12))
    pr, l1, l2 = pairing(12), g.standard_normal((12, 12)), g.standard_normal((12, 12))
    for args in ((), (pr, l1, l2)):
        f = ctx.ri_mo_fock_rkb(10, 11, nl, ns, hr, v, c, occ, h2, x2, *args)
        bmo = oracle.build_ri_mo_eri(nl, ns, bl.reshape(naux, -1), bs.reshape(naux, -1), v @ c)
        fr = oracle.ri_hx_fock_gemm(bmo, c.conj().T @ hr @ c, occ, h2, x2, *args)
        assert rel(f, fr) < TOL


def test_fused_functional_fock_nonrel(ctx):
    """riJkOnlyFockMatrixNonRel body (spatial half, embedded spin tensor, spin-blocked h_mo)."""
    from rerdmft_amd import synthetic as syn
    g = np.random.default_rng(410)
    ns, naux = 11, 45
    b = syn.ri_tensor(naux, ns, 0.5, 5.0, 411)
    ctx.upload(12, b)
    c = g.standard_normal((ns, 2 * ns))
    hc = syn.symmetric(ns, 412)
    occ = g.uniform(0, 1, 2 * ns)
    h2, x2 = g.standard_normal((2 * ns, 2 * ns)), g.standard_normal((2 * ns, 2 * ns))
    f = ctx.ri_mo_fock_nonrel(12, hc, c, occ, h2, x2)
    bmo = oracle.build_ri_nonrel_spin_mo_eri(b.reshape(naux, -1), c[:, :ns]).reshape(naux, ns, ns)
    emb = np.zeros((naux, 2 * ns, 2 * ns))
    emb[:, :ns, :ns] = bmo
    emb[:, ns:, ns:] = bmo
    hmo = c.T @ hc @ c
    hmo[:ns, ns:] = 0.0
    hmo[ns:, :ns] = 0.0
    fr = oracle.ri_hx_fock_gemm(emb, hmo, occ, h2, x2)
    assert rel(f, fr) < TOL
