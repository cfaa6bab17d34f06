"""The device DIIS against the oracle (oracle/diis.py, pinned in tests/test_oracle_diis.py): the
commutator error for real and complex matrices, and a sequence of extrapolations with a capped
history (weights and extrapolated matrices), the singular and zero-error cases, and the shape check."""
import numpy as np
import pytest

import rerdmft_amd as rr
from oracle import diis

pytestmark = pytest.mark.gpu


def _rand(g, shape, cplx):
    return g.standard_normal(shape) + (1j * g.standard_normal(shape) if cplx else 0)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [7, 64, 130])
def test_commutator_error_matches_oracle(ctx, cplx, n):
    g = np.random.default_rng(n + 10 * cplx)
    f, p, s = (_rand(g, (n, n), cplx) for _ in range(3))
    e = ctx.scf_commutator_error(f, p, s)
    ref = diis.commutator_error(f, p, s)
    assert np.max(np.abs(e - ref)) < 1e-13 * np.max(np.abs(ref)) * n


@pytest.mark.parametrize("cplx", [False, True])
def test_extrapolation_sequence_matches_oracle(ctx, cplx):
    g = np.random.default_rng(21 + cplx)
    n = 40
    d = ctx.diis_create(4)
    ref = diis.Diis(4)
    try:
        for k in range(9):
            e = _rand(g, (n, n), cplx) * 0.5 ** k
            v = _rand(g, (n, n), cplx)
            out, w = ctx.diis_extrapolate(d, e, v)
            out_ref = ref.extrapolate(e, v)
            assert len(w) == len(ref.last_weights)
            if len(w):
                np.testing.assert_allclose(w, ref.last_weights, rtol=1e-10, atol=1e-12)
            assert np.max(np.abs(out - out_ref)) < 1e-11 * np.max(np.abs(out_ref))
    finally:
        ctx.diis_destroy(d)


def test_singular_zero_and_disabled_cases(ctx):
    e = np.random.default_rng(5).standard_normal((6, 6))
    d = ctx.diis_create(5)
    try:
        for k in range(3):
            out, w = ctx.diis_extrapolate(d, e, np.full((6, 6), float(k)))
        assert len(w) == 0 and np.array_equal(out, np.full((6, 6), 2.0))
        ctx.diis_clear(d)
        for k in range(2):
            out, w = ctx.diis_extrapolate(d, np.zeros((6, 6)), np.full((6, 6), float(k)))
        assert len(w) == 0 and np.array_equal(out, np.full((6, 6), 1.0))
        with pytest.raises(rr.GpuError, match="shapes differ"):
            ctx.diis_extrapolate(d, np.zeros((5, 5)), np.zeros((5, 5)))
    finally:
        ctx.diis_destroy(d)
    d1 = ctx.diis_create(1)
    try:
        out, w = ctx.diis_extrapolate(d1, e, e + 1.0)
        assert len(w) == 0 and np.array_equal(out, e + 1.0)
    finally:
        ctx.diis_destroy(d1)
