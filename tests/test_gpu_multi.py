"""Many densities sharing one tensor and one left factor (the sigma-product trial densities
D_k = L R_k^T) in one graph-captured build: every J_k, K_k must equal the single-density factor build
and the oracle, on the eager first call and on graph replays."""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn

pytestmark = pytest.mark.gpu

SLOT = 12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _case(naux, n, r, nd, seed, symmetric=True):
    b = syn.ri_tensor(naux, n, 0.8, 8.0, seed, symmetric=symmetric)
    g = syn.rng(seed + 1)
    l = syn.orthonormal_occ(n, r, seed + 2)
    rs = g.standard_normal((nd, n, r))
    return b, l, rs


@pytest.mark.parametrize("shape", [(60, 45, 7, 5), (96, 130, 40, 11)])
def test_multi_matches_single_builds_and_oracle(ctx, oracle, shape):
    naux, n, r, nd = shape
    b, l, rs = _case(naux, n, r, nd, 4100 + n)
    ctx.upload(SLOT, b, rr.COPY_FP64)
    j, k, rep = ctx.ri_jk_multi(SLOT, l, rs)
    assert not rep
    j2, k2, rep2 = ctx.ri_jk_multi(SLOT, l, rs)
    assert rep2
    assert np.array_equal(j, j2) and np.array_equal(k, k2)
    for i in range(nd):
        js, ks = ctx.ri_jk_factor(SLOT, None, l, rs[i])
        assert rel(j[i], js) < 1e-13 and rel(k[i], ks) < 1e-13
        if i < 3:
            jo, ko = oracle.ri_jk_factor(b.reshape(naux, -1), l @ rs[i].T, l, rs[i])
            assert rel(j[i], jo) < 1e-12 and rel(k[i], ko) < 1e-12


def test_multi_replay_sees_new_inputs(ctx):
    naux, n, r, nd = 50, 40, 6, 4
    b, l, rs = _case(naux, n, r, nd, 4200)
    ctx.upload(SLOT, b, rr.COPY_FP64)
    ctx.ri_jk_multi(SLOT, l, rs)
    rs2 = rs[::-1].copy()
    l2 = syn.orthonormal_occ(n, r, 4201)
    j, k, rep = ctx.ri_jk_multi(SLOT, l2, rs2)
    assert rep
    for i in range(nd):
        js, ks = ctx.ri_jk_factor(SLOT, None, l2, rs2[i])
        assert rel(j[i], js) < 1e-13 and rel(k[i], ks) < 1e-13


@pytest.mark.parametrize("prec,tol", [(rr.FP32, 2e-5), (rr.BF16, 3e-2)])
def test_multi_low_precision(ctx, prec, tol):
    naux, n, r, nd = 80, 64, 12, 9
    b, l, rs = _case(naux, n, r, nd, 4300)
    ctx.upload(SLOT, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    j64, k64, _ = ctx.ri_jk_multi(SLOT, l, rs)
    j, k, _ = ctx.ri_jk_multi(SLOT, l, rs, precision=prec)
    for i in range(nd):
        assert rel(j[i], j64[i]) < tol and rel(k[i], k64[i]) < tol
        # same operands and accumulation as the single-density build: equal up to summation order
        js, ks = ctx.ri_jk_factor(SLOT, None, l, rs[i], precision=prec)
        assert rel(j[i], js) < 1e-12 and rel(k[i], ks) < 1e-12


def test_multi_nonsymmetric_tensor_and_partial_outputs(ctx):
    naux, n, r, nd = 30, 33, 5, 3
    b, l, rs = _case(naux, n, r, nd, 4400, symmetric=False)
    ctx.upload(SLOT, b, rr.COPY_FP64)
    j, _, _ = ctx.ri_jk_multi(SLOT, l, rs, want_k=False)
    _, k, _ = ctx.ri_jk_multi(SLOT, l, rs, want_j=False)
    for i in range(nd):
        js, ks = ctx.ri_jk_factor(SLOT, None, l, rs[i])
        assert rel(j[i], js) < 1e-13 and rel(k[i], ks) < 1e-13


def test_multi_rejects_mismatched_dimension(ctx):
    b, l, rs = _case(20, 16, 3, 2, 4500)
    ctx.upload(SLOT, b)
    with pytest.raises(rr.GpuError):
        ctx.ri_jk_multi(SLOT, np.zeros((17, 3)), np.zeros((2, 17, 3)))


@pytest.mark.parametrize("n", [72, 70])
def test_multi_stacked_rank_k_variants_agree_bitwise(ctx, n):
    """r = 60 puts 6 densities in a half-transform launch and stacks 12 of them in one rank-k launch
    (13 densities: two half-transforms per stack, then a partial stack of one). Every bf16 GEMM variant,
    the persistent rank-k kernel's column-block mode included (n = 72; at n = 70 its 4-column stores
    would straddle the blocks and the other kernels run), gives the same J and K bit for bit, equal to
    the single-density builds up to summation order."""
    naux, r, nd = 40, 60, 13
    b, l, rs = _case(naux, n, r, nd, 4400 + n)
    ctx.upload(SLOT, b, rr.COPY_BF16)
    outs = []
    try:
        for cfg in [-1, 1, 2, 5, 100, 101, 102, 112, 120, 125, 130, 134, 142]:
            ctx.set_tile_config(cfg)
            j, k, _ = ctx.ri_jk_multi(SLOT, l, rs, precision=rr.BF16)
            outs.append((cfg, j, k))
    finally:
        ctx.set_tile_config(-1)
    for cfg, j, k in outs[1:]:
        assert np.array_equal(k, outs[0][2]), f"K differs for configuration {cfg}"
        assert np.array_equal(j, outs[0][1]), f"J differs for configuration {cfg}"
    for i in range(nd):
        js, ks = ctx.ri_jk_factor(SLOT, None, l, rs[i], precision=rr.BF16)
        assert rel(outs[0][1][i], js) < 1e-12 and rel(outs[0][2][i], ks) < 1e-12
