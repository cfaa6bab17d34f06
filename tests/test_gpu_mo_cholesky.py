"""GPU parity of the MO-basis RI tensors (buildRiMoEri, rotations, NON_REL spin tensors) and of the
Cholesky Fock builds / Cholesky MO vectors against the oracle, plus size-independent properties
(rotation round trip, unitary invariance)."""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn
from tests._lcg import Lcg
from tests.test_oracle import lcg_ao_inputs

pytestmark = pytest.mark.gpu
TOL = {rr.FP64: 1e-11, rr.FP32: 3e-6, rr.BF16: 3e-2}


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32, rr.BF16])
@pytest.mark.parametrize("ns", [0, 13])
def test_build_ri_mo_eri(ctx, oracle, prec, ns):
    nl, naux, nt = 21, 37, 30
    bll = syn.ri_tensor(naux, nl, 0.5, 5.0, 701)
    bss = syn.ri_tensor(naux, max(ns, 1), 0.3, 5.0, 702)[:, :ns, :ns] if ns else np.zeros((naux, 0, 0))
    n = 2 * nl + 2 * ns
    g = syn.rng(703)
    c = g.standard_normal((n, nt)) + 1j * g.standard_normal((n, nt))
    ctx.upload(0, bll, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    if ns:
        ctx.upload(1, bss, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    got = ctx.build_ri_mo_eri(0, 1, nl, ns, c, precision=prec, out_precision=rr.FP64 if prec == rr.FP64 else rr.FP32)
    ref = oracle.build_ri_mo_eri(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), c)
    assert rel(got.astype(np.complex128), ref) < TOL[prec]


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32])
def test_ri_mo_eri_rotation(ctx, oracle, prec):
    naux, nt = 23, 19
    g = syn.rng(711)
    b = g.standard_normal((naux, nt, nt)) + 1j * g.standard_normal((naux, nt, nt))
    u = syn.unitary(nt, 712)
    got = ctx.ri_mo_eri_rotated(b, u, precision=prec)
    ref = oracle.ri_mo_eri_rotated(b, u)
    assert rel(got, ref) < TOL[prec]
    if prec == rr.FP64:
        back = ctx.ri_mo_eri_rotated(got, u.conj().T)
        assert rel(back, b) < 1e-12


def test_positive_block(ctx):
    g = syn.rng(721)
    b = g.standard_normal((5, 9, 9)) + 1j * g.standard_normal((5, 9, 9))
    got = ctx.ri_mo_eri_positive_block(b, 4)
    assert np.array_equal(got, b[:, 4:, 4:])


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32, rr.BF16])
def test_nonrel_spin_mo_eri_and_rotation(ctx, oracle, prec):
    naux, n = 31, 26
    b = syn.ri_tensor(naux, n, 0.5, 6.0, 731)
    c = syn.orthonormal_occ(n, n, 732)
    ctx.upload(2, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    got = ctx.build_ri_nonrel_spin_mo_eri(2, c, precision=prec)
    ref = oracle.build_ri_nonrel_spin_mo_eri(b.reshape(naux, -1), c)
    assert rel(got, ref) < TOL[prec]
    us = syn.orthonormal_occ(n, n, 733)
    u = np.kron(np.eye(2), us)
    rot = ctx.ri_nonrel_spin_mo_eri_rotated(ref, u, precision=prec)
    rref = oracle.ri_nonrel_spin_mo_eri_rotated(ref, u)
    assert rel(rot, rref) < TOL[prec]
    emb = ctx.ri_nonrel_spin_embedded_b(ref)
    assert np.array_equal(emb[:, :n, :n], ref) and np.array_equal(emb[:, n:, n:], ref)
    assert not np.any(emb[:, :n, n:]) and not np.any(emb[:, n:, :n])


def test_nonrel_spin_rotation_rejects_unrestricted(ctx):
    b = np.zeros((2, 3, 3))
    u = np.eye(6)
    u[0, 3] = 0.1
    with pytest.raises(rr.GpuError, match="not spin-restricted"):
        ctx.ri_nonrel_spin_mo_eri_rotated(b, u)


def test_mo_transform_rectangular_nmo(ctx, oracle):
    # n_mo != n (occupied-only transform) at a ragged size
    naux, n, nmo = 17, 45, 11
    b = syn.ri_tensor(naux, n, 0.5, 8.0, 741)
    c = syn.rng(742).standard_normal((n, nmo))
    ctx.upload(3, b)
    got = ctx.build_ri_nonrel_spin_mo_eri(3, c)
    ref = np.einsum("mp,Pmn,nq->Ppq", c, b, c)
    assert rel(got, ref) < 1e-12


# ---- Cholesky ---------------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32])
def test_cholesky_fock_on_reference_inputs(ctx, oracle, prec):
    vec, h, p, hx, px, _ = lcg_ao_inputs()
    nchol = vec.shape[0]
    ctx.upload(4, vec, rr.COPY_FP64 | rr.COPY_FP32)
    f = ctx.nonrel_fock_cholesky(4, h, p, precision=prec)
    fr = oracle.nonrel_fock_cholesky(vec.reshape(nchol, -1), h, p)
    assert rel(f, fr) < TOL[prec]
    fx = ctx.x2c_fock_cholesky(4, hx, px, precision=prec)
    fxr = oracle.x2c_fock_cholesky(vec.reshape(nchol, -1), hx, px)
    assert rel(fx, fxr) < TOL[prec]


def test_cholesky_fock_larger(ctx, oracle):
    n, nchol = 70, 140
    vec = syn.ri_tensor(nchol, n, 0.3, 10.0, 751)
    h = syn.symmetric(n, 752)
    p = syn.symmetric(n, 753)
    ctx.upload(5, vec)
    assert rel(ctx.nonrel_fock_cholesky(5, h, p), oracle.nonrel_fock_cholesky(vec.reshape(nchol, -1), h, p)) < 1e-11
    hx = syn.hermitian(2 * n, 754)
    px = syn.hermitian(2 * n, 755)
    assert rel(ctx.x2c_fock_cholesky(5, hx, px), oracle.x2c_fock_cholesky(vec.reshape(nchol, -1), hx, px)) < 1e-11


def test_cholesky_mo_vectors(ctx, oracle):
    vec, _, _, _, _, rng = lcg_ao_inputs()
    nchol, n, nmo = vec.shape[0], vec.shape[1], 5
    cr = np.array([[rng.next() for _ in range(nmo)] for _ in range(n)])
    ctx.upload(6, vec)
    w = ctx.cholesky_to_mo_spatial(6, cr)
    assert rel(w, oracle.cholesky_to_mo_spatial(vec.reshape(nchol, -1), cr)) < 1e-12
    cc = np.array([[complex(rng.next(), rng.next()) for _ in range(nmo)] for _ in range(2 * n)])
    ws = ctx.cholesky_to_mo_spinor(6, cc)
    assert rel(ws, oracle.cholesky_to_mo_spinor(vec.reshape(nchol, -1), cc)) < 1e-12


@pytest.mark.parametrize("prec,tol", [(rr.FP64, 1e-12), (rr.FP32, 2e-5), (rr.BF16, 3e-2)])
def test_cholesky_factor_form_equals_density_form(ctx, oracle, prec, tol):
    # AoCholesky.cpp:133-136 builds K = sum_L B_L P B_L from the density; for P = 2 C C^T the same K is
    # sum_L (B_L sqrt2 C)(B_L sqrt2 C)^T, the factor-form exchange with N_chol in place of n_aux
    n, nchol, nocc = 64, 150, 9
    vec = syn.ri_tensor(nchol, n, 0.3, 10.0, 761)
    c = syn.orthonormal_occ(n, nocc, 762)
    p = 2.0 * c @ c.T
    h = syn.symmetric(n, 763)
    ctx.upload(7, vec, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    f_ref = oracle.nonrel_fock_cholesky(vec.reshape(nchol, -1), h, p)
    assert rel(ctx.nonrel_fock_cholesky(7, h, p), f_ref) < 1e-12
    j, k = ctx.ri_jk_factor(7, None, np.sqrt(2.0) * c, precision=prec)
    assert rel(h + j - 0.5 * k, f_ref) < tol


@pytest.mark.parametrize("prec,tol", [(rr.FP64, 1e-11), (rr.FP32, 2e-5), (rr.BF16, 3e-2)])
def test_cholesky_factor_form_from_a_density(ctx, oracle, prec, tol):
    """Factor-form Cholesky builds from a density: the RI builds' occupied-set recovery
    (NonRelFockMatrixRi.cpp:51-69, X2C_FockMatrixRi.cpp:65-95) run on a slot of Cholesky vectors give
    K = sum_L (B_L C)(B_L C)^H with N_chol in place of n_aux, equal to the density form
    K = sum_L B_L P B_L of AoCholesky.cpp:108-217 for an SCF density (P = C occ C^H, C^H S C = I)."""
    from tests.test_gpu_relativistic import x2c_inputs
    nl, nchol = 30, 110
    b, s, x, p, h = x2c_inputs(nl, nchol, 6, 771)
    ctx.upload(8, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    f_ref = oracle.x2c_fock_cholesky(b.reshape(nchol, -1), h, p)
    assert rel(ctx.x2c_fock_cholesky(8, h, p), f_ref) < 1e-11
    assert rel(ctx.x2c_fock_ri(8, h, x, s, p, precision=prec), f_ref) < tol
    # NON_REL on the alpha block's spatial problem
    s1, x1 = s[:nl, :nl].real.copy(), x[:nl, :nl].real.copy()
    c = x1 @ syn.orthonormal_occ(nl, 5, 772)
    pn = 2.0 * c @ c.T
    hn = syn.symmetric(nl, 773)
    g_ref = oracle.nonrel_fock_cholesky(b.reshape(nchol, -1), hn, pn)
    assert rel(ctx.nonrel_fock_ri(8, hn, x1, s1, pn, precision=prec), g_ref) < tol
