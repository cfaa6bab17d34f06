"""GPU parity of the NON_REL RI J/K path (librerdmft_gpu through its C ABI) against the oracle.

Tolerances: FP64 mode must match the FP64 oracle to 1e-12 relative to max|K| (tighter than the
1e-10 Ha energy bar); FP32 operands with FP32 accumulation to 2e-6 relative; BF16 operands with
FP32 accumulation to 2e-2 relative (bf16 keeps 8 significand bits).
"""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn

pytestmark = pytest.mark.gpu

TOL = {rr.FP64: 1e-12, rr.FP32: 2e-6, rr.BF16: 2e-2}


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def water_like(seed=syn.BASE_SEED + 0):
    """Kernel-level water/cc-pVDZ case: B[300][24][24], N(0,0.1) decaying exp(-|mu-nu|/8), 5 occ."""
    b = syn.ri_tensor(300, 24, 0.1, 8.0, seed)
    c = syn.orthonormal_occ(24, 5, seed + 100)
    return b, c


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32, rr.BF16])
def test_jk_factor_water_like(ctx, oracle, prec):
    b, c = water_like()
    d = syn.density(c)
    l = np.sqrt(2.0) * c
    ctx.upload(0, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    j, k = ctx.ri_jk_factor(0, d, l, precision=prec)
    jr, kr = oracle.ri_jk_factor(b.reshape(300, -1), d, l)
    assert rel(j, jr) < TOL[prec]
    assert rel(k, kr) < TOL[prec]


@pytest.mark.parametrize("shape", [(37, 33, 7), (200, 130, 29), (64, 96, 200), (17, 5, 1)])
def test_jk_factor_ragged_shapes(ctx, oracle, shape):
    naux, n, r = shape
    g = syn.rng(naux * 1000 + n)
    b = syn.ri_tensor(naux, n, 1.0, 10.0, naux + n)
    b[:, 0, 1] += 0.3  # non-symmetric rows pin the contraction index order
    d = syn.symmetric(n, 5)
    d[2, 3] += 0.7     # non-symmetric density pins the Coulomb index order
    l = g.standard_normal((n, r))
    rf = g.standard_normal((n, r))
    ctx.upload(1, b, rr.COPY_FP64)
    j, k = ctx.ri_jk_factor(1, d, l, precision=rr.FP64)
    jr, kr = oracle.ri_jk_factor(b.reshape(naux, -1), d, l)
    assert rel(j, jr) < 1e-12 and rel(k, kr) < 1e-12
    j2, k2 = ctx.ri_jk_factor(1, d, l, rf, precision=rr.FP64)
    _, kr2 = oracle.ri_jk_factor(b.reshape(naux, -1), d, l, rf)
    assert rel(k2, kr2) < 1e-12


def test_nonrel_fock_ri_with_recovery(ctx, oracle):
    n, naux, nocc = 40, 150, 9
    b = syn.ri_tensor(naux, n, 0.3, 6.0, 77)
    s = syn.spd_overlap(n, 78)
    x = syn.inverse_sqrt(s)
    c = syn.s_orthonormal_occ(s, nocc, 79)
    p = syn.density(c)
    h = syn.symmetric(n, 80)
    ctx.upload(2, b, rr.COPY_FP64 | rr.COPY_FP32)
    f, j, k = ctx.nonrel_fock_ri(2, h, x, s, p, return_jk=True)
    fr, jr, kr = oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p, return_jk=True)
    assert rel(f, fr) < 1e-12 and rel(j, jr) < 1e-12 and rel(k, kr) < 1e-12
    e_gpu = 0.5 * np.sum(p * (h + f))
    e_ref = 0.5 * np.sum(p * (h + fr))
    assert abs(e_gpu - e_ref) < 1e-10
    f32 = ctx.nonrel_fock_ri(2, h, x, s, p, precision=rr.FP32)
    assert rel(f32, fr) < 1e-5


def test_nonrel_fock_ri_non_idempotent_threshold(ctx, oracle):
    n, naux = 20, 60
    b = syn.ri_tensor(naux, n, 0.3, 6.0, 91)
    s = syn.spd_overlap(n, 92)
    x = syn.inverse_sqrt(s)
    c = syn.s_orthonormal_occ(s, 3, 93)
    p = 2.0 * (0.95 * np.outer(c[:, 0], c[:, 0]) + 0.6 * np.outer(c[:, 1], c[:, 1]) + 0.2 * np.outer(c[:, 2], c[:, 2]))
    h = syn.symmetric(n, 94)
    ctx.upload(3, b)
    f = ctx.nonrel_fock_ri(3, h, x, s, p)
    fr = oracle.nonrel_fock_ri(b.reshape(naux, -1), x, s, h, p)
    assert rel(f, fr) < 1e-12


def test_nonrel_fock_ri_dimension_error_matches_reference_message(ctx):
    b = syn.ri_tensor(10, 6, 0.3, 3.0, 1)
    ctx.upload(4, b)
    h = np.zeros((5, 5))
    with pytest.raises(rr.GpuError, match="density_matrix dimension does not match nl"):
        ctx.nonrel_fock_ri(4, h, np.eye(5), np.eye(5), np.zeros((5, 5)))


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32, rr.BF16])
def test_delta_build_matches_difference_of_full_builds(ctx, oracle, prec):
    n, naux, nocc = 96, 120, 12
    b = syn.ri_tensor(naux, n, 1.0, 20.0, 101)
    q, _ = np.linalg.qr(syn.rng(102).standard_normal((n, n)))
    c_old = np.ascontiguousarray(q[:, :nocc])
    c_new = syn.rotate_occ(c_old, q, 0.05, 103)
    # scramble the gauge of c_new inside its occupied space: the density is unchanged
    u = syn.orthonormal_occ(nocc, nocc, 104)
    c_new_g = c_new @ u
    ctx.upload(5, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    dj, dk = ctx.ri_jk_delta(5, c_new_g, c_old, precision=prec)
    jn, kn = oracle.ri_jk_factor(b.reshape(naux, -1), syn.density(c_new), np.sqrt(2) * c_new)
    jo, ko = oracle.ri_jk_factor(b.reshape(naux, -1), syn.density(c_old), np.sqrt(2) * c_old)
    tol = {rr.FP64: 1e-10, rr.FP32: 1e-5, rr.BF16: 3e-2}[prec]
    assert rel(dj, jn - jo) < tol
    assert rel(dk, kn - ko) < tol
    # the error is relative to the change, not to the full matrices: check against |K|
    if prec == rr.BF16:
        assert np.max(np.abs(dk - (kn - ko))) < 1e-3 * np.max(np.abs(kn))


def test_tile_configurations_agree_bitwise(ctx):
    b, c = water_like(syn.BASE_SEED + 5)
    ctx.upload(6, b, rr.COPY_FP64 | rr.COPY_FP32)
    d = syn.density(c)
    l = np.sqrt(2.0) * c
    try:
        for prec in (rr.FP64, rr.FP32):
            outs = []
            for cfg in list(range(8)) + list(range(100, 107)) + list(range(126, 130)) + [137, 138]:
                ctx.set_tile_config(cfg)
                outs.append(ctx.ri_jk_factor(6, d, l, precision=prec)[1])
            ctx.set_tile_config(-1)
            outs.append(ctx.ri_jk_factor(6, d, l, precision=prec)[1])
            for o in outs[1:]:
                assert np.array_equal(o, outs[0])
    finally:
        ctx.set_tile_config(-1)


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32, rr.BF16])
def test_fused_coulomb_from_factors(ctx, oracle, prec):
    # density=None: J of D = L R^T taken from the half-transformed factors inside the exchange GEMM
    naux, n, r = 150, 70, 13
    b = syn.ri_tensor(naux, n, 1.0, 10.0, 121)
    b[:, 3, 5] += 0.2
    g = syn.rng(122)
    l, rf = g.standard_normal((n, r)), g.standard_normal((n, r))
    ctx.upload(8, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    j, k = ctx.ri_jk_factor(8, None, l, precision=prec)
    jr, kr = oracle.ri_jk_factor(b.reshape(naux, -1), l @ l.T, l)
    tol = {rr.FP64: 1e-12, rr.FP32: 2e-6, rr.BF16: 2e-2}[prec]
    assert rel(j, jr) < tol and rel(k, kr) < tol
    j2, k2 = ctx.ri_jk_factor(8, None, l, rf, precision=prec)
    jr2, kr2 = oracle.ri_jk_factor(b.reshape(naux, -1), l @ rf.T, l, rf)
    assert rel(j2, jr2) < tol and rel(k2, kr2) < tol


def test_empty_occupied_space(ctx):
    b = syn.ri_tensor(8, 6, 0.3, 3.0, 2)
    ctx.upload(7, b)
    j, k = ctx.ri_jk_factor(7, np.zeros((6, 6)), np.zeros((6, 0)))
    assert np.all(k == 0.0) and np.all(j == 0.0)


@pytest.mark.parametrize("naux,n,r", [(40, 37, 7), (120, 130, 29), (64, 200, 64), (24, 330, 100)])
def test_bf16_gemm_variants_agree_bitwise(ctx, naux, n, r):
    """Every bf16 GEMM variant (register-staged, LDS-staged single/double buffered, multi-stage
    direct-to-LDS, persistent row-merged and split-K) sums each output in the same k order: results must be
    bitwise identical."""
    b = syn.ri_tensor(naux, n, 1.0, 10.0, 300 + n)
    g = syn.rng(301 + n)
    c_old = syn.orthonormal_occ(n, r, 302 + n)
    q, _ = np.linalg.qr(g.standard_normal((n, n)))
    c_new = syn.rotate_occ(c_old, q, 0.05, 303 + n)
    l = g.standard_normal((n, r))
    ctx.upload(9, b, rr.COPY_BF16)
    outs = []
    try:
        for cfg in [1, 2, 5, 100, 101] + list(range(102, 126)) + list(range(130, 137)) + [139, 140, 141, 142]:
            ctx.set_tile_config(cfg)
            dj, dk = ctx.ri_jk_delta(9, c_new, c_old, precision=rr.BF16)
            j, k = ctx.ri_jk_factor(9, None, l, precision=rr.BF16)
            outs.append((cfg, dk, k, dj, j))
    finally:
        ctx.set_tile_config(-1)
    for cfg, dk, k, dj, j in outs[1:]:
        assert np.array_equal(dk, outs[0][1]), f"incremental K differs for configuration {cfg}"
        assert np.array_equal(k, outs[0][2]), f"K differs for configuration {cfg}"
        # the Coulomb contraction fused into the half-transform epilogue is tile-independent too
        assert np.array_equal(dj, outs[0][3]), f"incremental J differs for configuration {cfg}"
        assert np.array_equal(j, outs[0][4]), f"J differs for configuration {cfg}"


@pytest.mark.parametrize("symmetric", [True, False])
@pytest.mark.parametrize("n", [37, 130])
def test_coulomb_upper_triangle_and_full_stream(ctx, symmetric, n):
    """An exactly symmetric tensor is streamed over its upper triangle only; a non-symmetric one over
    every element. Both must give J = sum_P tr(B_P D) B_P (D symmetric), in every precision."""
    naux = 23
    b = syn.ri_tensor(naux, n, 0.7, 6.0, 900 + n, symmetric=symmetric)
    d = syn.symmetric(n, 901 + n)
    gamma = np.einsum("pmn,nm->p", b, d)
    ref = np.einsum("p,pmn->mn", gamma, b)
    ctx.upload(6, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    for prec, tol in [(rr.FP64, 1e-12), (rr.FP32, 1e-5), (rr.BF16, 3e-2)]:
        j, _ = ctx.ri_jk_factor(6, d, None, precision=prec, want_k=False)
        assert np.max(np.abs(j - ref)) <= tol * np.max(np.abs(ref)), (prec, symmetric)
        if prec == rr.FP64 and symmetric:
            assert np.array_equal(j, j.T)


@pytest.mark.parametrize("naux,n,r", [(40, 37, 7), (64, 200, 64), (30, 300, 150)])
def test_fp64_gemm_variants_agree_bitwise(ctx, naux, n, r):
    """The eight-wave fp64 direct-to-LDS kernels take the same matrix-core steps in the same k order as
    the register-direct and LDS-staged fp64 kernels: J and K must be bitwise identical."""
    b = syn.ri_tensor(naux, n, 1.0, 10.0, 500 + n)
    g = syn.rng(501 + n)
    l, rf = g.standard_normal((n, r)), g.standard_normal((n, r))
    ctx.upload(9, b, rr.COPY_FP64)
    outs = []
    try:
        for cfg in [1, 2, 5, 100, 101, 126, 127, 128, 129, 137, 138]:
            ctx.set_tile_config(cfg)
            j, k = ctx.ri_jk_factor(9, None, l)          # symmetric K: triangular rank-k update
            j2, k2 = ctx.ri_jk_factor(9, None, l, rf)    # full rank-k update
            outs.append((cfg, j, k, j2, k2))
    finally:
        ctx.set_tile_config(-1)
    for cfg, *res in outs[1:]:
        for a, b0 in zip(res, outs[0][1:]):
            assert np.array_equal(a, b0), f"fp64 result differs for configuration {cfg}"
