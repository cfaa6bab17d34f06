"""CPU pins of the DIIS oracle (oracle/diis.py): the commutator error vanishes at self-consistency
(and a planted P S F is caught), is antisymmetric for symmetric inputs; the extrapolation weights
match the two-vector closed form, a linear residual is extrapolated to its root exactly, and a
singular history returns the value unchanged."""
import numpy as np
import pytest
import scipy.linalg

from oracle import diis


def _consistent(n, nocc, seed, cplx=False):
    g = np.random.default_rng(seed)
    a = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0)
    s = (a @ a.conj().T) / n + np.eye(n)
    f = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0)
    f = f + f.conj().T
    w, c = scipy.linalg.eigh(f, s)
    p = c[:, :nocc] @ c[:, :nocc].conj().T
    return f, p, s


@pytest.mark.parametrize("cplx", [False, True])
def test_commutator_vanishes_at_self_consistency(cplx):
    f, p, s = _consistent(12, 4, 1, cplx)
    assert np.max(np.abs(diis.commutator_error(f, p, s))) < 1e-11 * np.max(np.abs(f))
    # the planted P S F form does not vanish
    assert np.max(np.abs(diis.commutator_error(f, p, s, planted=True))) > 1e-3


def test_commutator_is_antisymmetric_for_symmetric_inputs():
    g = np.random.default_rng(2)
    f, p, s = (x + x.T for x in (g.standard_normal((9, 9)) for _ in range(3)))
    e = diis.commutator_error(f, p, s)
    np.testing.assert_allclose(e, -e.T, atol=1e-12)
    assert np.max(np.abs(e)) > 1e-2


def test_two_vector_weights_match_closed_form():
    g = np.random.default_rng(3)
    e1, e2 = g.standard_normal(20), g.standard_normal(20)
    v1, v2 = g.standard_normal(7), g.standard_normal(7)
    d = diis.Diis(5)
    out0 = d.extrapolate(e1, v1)
    np.testing.assert_array_equal(out0, v1)  # one pair: returned unchanged
    assert d.last_weights == []
    out = d.extrapolate(e2, v2)  # newest first: weights (w_new, w_old)
    b11, b12, b22 = e2 @ e2, e2 @ e1, e1 @ e1
    w_new = (b22 - b12) / (b11 - 2 * b12 + b22)
    np.testing.assert_allclose(d.last_weights, [w_new, 1 - w_new], rtol=1e-12)
    np.testing.assert_allclose(out, w_new * v2 + (1 - w_new) * v1, rtol=1e-12)


@pytest.mark.parametrize("cplx", [False, True])
def test_linear_residual_is_extrapolated_to_its_root(cplx):
    g = np.random.default_rng(4)
    shape = (6, 6)
    root = g.standard_normal(shape) + (1j * g.standard_normal(shape) if cplx else 0)
    basis = [g.standard_normal(shape) + (1j * g.standard_normal(shape) if cplx else 0) for _ in range(2)]
    d = diis.Diis(4)
    # three values whose residuals V_i - root lie in a 2-dimensional space: the combination with
    # weights summing to one that cancels the residual reproduces the root
    coefs = [(1.0, 0.3), (-0.4, 0.8), (0.2, -1.1)]
    out = None
    for a, b in coefs:
        v = root + a * basis[0] + b * basis[1]
        out = d.extrapolate(v - root, v)
    np.testing.assert_allclose(out, root, atol=1e-10)


def test_identical_errors_are_singular_and_return_the_value():
    e = np.random.default_rng(5).standard_normal(10)
    d = diis.Diis(5)
    for k in range(3):
        out = d.extrapolate(e, np.full(4, float(k)))
    np.testing.assert_array_equal(out, np.full(4, 2.0))
    assert d.last_weights == []


def test_history_is_capped_newest_first():
    g = np.random.default_rng(6)
    d = diis.Diis(3)
    for k in range(6):
        d.extrapolate(g.standard_normal(8), np.full(2, float(k)))
    assert len(d.errors) == 3 and [v[0] for v in d.values] == [5.0, 4.0, 3.0]
