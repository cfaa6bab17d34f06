"""Batches of independent molecules in one HIP-graph launch: every molecule's J and K must equal the
single-molecule rr_ri_jk_factor result (same kernels, same tiles: bitwise), on the first (eager +
capture) call and on every replay, with eager single-molecule work and synchronizations in between."""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn
from oracle import oracle

pytestmark = pytest.mark.gpu


def molecules(count, seed):
    g = np.random.default_rng(seed)
    ns = g.integers(10, 60, count)
    mols = []
    for i, n in enumerate(ns):
        n = int(n)
        b = syn.ri_tensor(4 * n, n, 0.3, 5.0, seed + i)
        mols.append((b, max(1, n // 5)))
    return mols


def factors(mols, seed):
    g = np.random.default_rng(seed)
    return [g.standard_normal((b.shape[1], r)) for b, r in mols]


@pytest.mark.parametrize("prec", [rr.FP64, rr.BF16])
def test_batch_matches_single_builds_and_replays(ctx, prec):
    mols = molecules(24, 500 + prec)
    copies = rr.COPY_FP64 | rr.COPY_BF16
    batch = ctx.batch_create([b for b, _ in mols], copies)
    try:
        # single-molecule builds first, so that they allocate their own input/output buffers once
        # (a workspace reallocation after a capture forces the next batch call to capture again)
        for (b, _), l in zip(mols, factors(mols, 599)):
            ctx.upload(13, b, copies)
            ctx.ri_jk_factor(13, None, l, precision=prec)
        for call in range(4):
            ls = factors(mols, 600 + call)
            js, ks, replayed = ctx.batch_jk_factor(batch, ls, precision=prec)
            assert replayed == (call > 0)
            for (b, _), l, j, k in zip(mols, ls, js, ks):
                ctx.upload(13, b, copies)  # eager single-molecule work between replays
                j1, k1 = ctx.ri_jk_factor(13, None, l, precision=prec)
                assert np.array_equal(j, j1) and np.array_equal(k, k1)
            if prec == rr.FP64 and call in (0, 2):
                b, _ = mols[call]
                jr, kr = oracle.ri_jk_factor(b.reshape(b.shape[0], -1), ls[call] @ ls[call].T, ls[call])
                assert np.max(np.abs(js[call] - jr)) < 1e-11 * np.max(np.abs(jr))
                assert np.max(np.abs(ks[call] - kr)) < 1e-11 * np.max(np.abs(kr))
    finally:
        ctx.batch_destroy(batch)


def test_batch_rebuilds_graph_when_ranks_change(ctx):
    mols = molecules(6, 700)
    batch = ctx.batch_create([b for b, _ in mols])
    try:
        ls = factors(mols, 701)
        ctx.batch_jk_factor(batch, ls)
        _, _, rep = ctx.batch_jk_factor(batch, ls)
        assert rep
        ls2 = [l[:, : max(1, l.shape[1] - 1)] for l in ls]
        js, ks, rep = ctx.batch_jk_factor(batch, ls2)
        assert not rep
        b, _ = mols[0]
        jr, kr = oracle.ri_jk_factor(b.reshape(b.shape[0], -1), ls2[0] @ ls2[0].T, ls2[0])
        assert np.max(np.abs(ks[0] - kr)) < 1e-11 * np.max(np.abs(kr))
    finally:
        ctx.batch_destroy(batch)


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32])
def test_batch_ragged_sizes_and_empty_factor(ctx, prec):
    # tiny and ragged orbital counts, and a molecule with no occupied orbitals (J = K = 0)
    sizes, ranks = [3, 8, 17, 33, 12], [1, 0, 4, 7, 12]
    mols = [syn.ri_tensor(max(1, 4 * n - 5), n, 0.3, 5.0, 900 + n) for n in sizes]
    copies = rr.COPY_FP64 | rr.COPY_FP32
    batch = ctx.batch_create(mols, copies)
    try:
        g = np.random.default_rng(901)
        ls = [g.standard_normal((n, r)) for n, r in zip(sizes, ranks)]
        for call in range(2):
            js, ks, rep = ctx.batch_jk_factor(batch, ls, precision=prec)
            assert rep == (call > 0)
            for b, l, j, k in zip(mols, ls, js, ks):
                if l.shape[1] == 0:
                    assert not np.any(j) and not np.any(k)
                    continue
                ctx.upload(13, b, copies)
                j1, k1 = ctx.ri_jk_factor(13, None, l, precision=prec)
                assert np.array_equal(j, j1) and np.array_equal(k, k1)
    finally:
        ctx.batch_destroy(batch)
