"""CASSCF RI contractions on the device against the oracle (oracle/cas.py, pinned in
tests/test_oracle_cas.py against the dense four-index tensor): the two-electron operator for real
and complex MO tensors with definite, indefinite and low-rank densities, the active integrals and
the active cumulant Fock matrix, plus a tensor built by the device MO transform."""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn
from oracle import cas

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _tensor(naux, n, seed, cplx):
    g = np.random.default_rng(seed)
    b = g.standard_normal((naux, n, n)) + (1j * g.standard_normal((naux, n, n)) if cplx else 0)
    return 0.5 * (b + b.conj().transpose(0, 2, 1))


def _rho(n, seed, cplx, rank=None):
    g = np.random.default_rng(seed)
    a = g.standard_normal((n, n)) + (1j * g.standard_normal((n, n)) if cplx else 0)
    r = a + a.conj().T
    if rank is not None:
        w, u = np.linalg.eigh(r)
        r = (u[:, :rank] * w[:rank]) @ u[:, :rank].conj().T
    return r


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape,rank", [((9, 13), None), ((40, 70), 6), ((123, 37), None)])
def test_two_electron_operator(ctx, cplx, shape, rank):
    naux, n = shape
    b = _tensor(naux, n, n + naux, cplx)
    rho = _rho(n, n, cplx, rank)
    got = ctx.cas_two_electron_operator(b, rho)
    ref = cas.two_electron_operator(b, n, rho)
    assert rel(got, ref) < 1e-12


def test_two_electron_operator_zero_density(ctx):
    b = _tensor(5, 6, 1, False)
    g = ctx.cas_two_electron_operator(b, np.zeros((6, 6)))
    assert np.all(g == 0.0)


@pytest.mark.parametrize("cplx", [False, True])
def test_active_integrals_and_cumulant_fock(ctx, cplx):
    naux, n = 50, 24
    b = _tensor(naux, n, 7, cplx)
    act = [2, 3, 9, 17, 20]
    a = len(act)
    np.testing.assert_allclose(ctx.cas_active_integrals(b, n, act), cas.active_integrals(b, n, act), rtol=0,
                               atol=1e-11 * np.max(np.abs(b)) ** 2 * naux)
    g = np.random.default_rng(8)
    lam = (g.standard_normal(a ** 4) + (1j * g.standard_normal(a ** 4) if cplx else 0))
    got = ctx.cas_active_cumulant_fock(b, n, act, lam)
    ref = cas.active_cumulant_fock(b, n, act, lam)
    assert rel(got, ref) < 1e-12


def test_on_device_mo_tensor(ctx):
    # the RI MO tensor of buildRiMoEri (Hermitian blocks by construction) feeding the CAS operator
    nl, naux, nt = 12, 30, 24
    bll = syn.ri_tensor(naux, nl, 0.5, 5.0, 771)
    c = np.linalg.qr(syn.rng(772).standard_normal((2 * nl, nt)) + 1j * syn.rng(773).standard_normal((2 * nl, nt)))[0]
    ctx.upload(0, bll)
    bmo = ctx.build_ri_mo_eri(0, 0, nl, 0, c)
    assert np.max(np.abs(bmo - bmo.conj().transpose(0, 2, 1))) < 1e-12 * np.max(np.abs(bmo))
    rho = _rho(nt, 774, True, 5)
    assert rel(ctx.cas_two_electron_operator(bmo, rho), cas.two_electron_operator(bmo, nt, rho)) < 1e-12


def test_rejects_bad_active_index(ctx):
    b = _tensor(4, 5, 9, False)
    with pytest.raises(rr.GpuError, match="active index out of range"):
        ctx.cas_active_integrals(b, 5, [0, 5])


@pytest.mark.parametrize("cplx", [False, True])
def test_cc_physics_block(ctx, cplx):
    naux, n = 40, 30
    b = _tensor(naux, n, 13, cplx)
    p, q, r, s = [0, 7, 29], list(range(5, 20)), [3, 4], list(range(30))
    got = ctx.cc_physics_block(b, n, p, q, r, s)
    ref = cas.physics_block(b, n, p, q, r, s)
    assert got.shape == (3, 15, 2, 30)
    assert rel(got, ref) < 1e-12
    with pytest.raises(rr.GpuError, match="index out of range"):
        ctx.cc_physics_block(b, n, [30], q, r, s)
