"""GPU parity of the complex RI paths (X2C 2-spinor and 4c RKB/UKB blocks) against the oracle.

FP64: 1e-11 relative to max|F|; FP32: 3e-6; BF16: 3e-2 (bf16 operands, fp32 accumulation).
"""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn

pytestmark = pytest.mark.gpu
TOL = {rr.FP64: 1e-11, rr.FP32: 3e-6, rr.BF16: 3e-2}


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def x2c_inputs(nl, naux, npairs, seed):
    b = syn.ri_tensor(naux, nl, 0.5, nl / 4.0, seed)
    s1 = syn.spd_overlap(nl, seed + 1)
    s = np.kron(np.eye(2), s1).astype(np.complex128)
    x = np.kron(np.eye(2), syn.inverse_sqrt(s1)).astype(np.complex128)
    q = syn.kramers_pair_occ(nl, npairs, seed + 2)
    c = x @ q
    p = c @ c.conj().T
    h = syn.hermitian(2 * nl, seed + 3)
    return b, s, x, p, h


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32, rr.BF16])
def test_x2c_fock_ri(ctx, oracle, prec):
    nl, naux = 36, 90
    b, s, x, p, h = x2c_inputs(nl, naux, 5, 301)
    ctx.upload(0, b, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    f = ctx.x2c_fock_ri(0, h, x, s, p, precision=prec)
    fr = oracle.x2c_fock_ri(b.reshape(naux, -1), x, s, h, p)
    assert rel(f, fr) < TOL[prec]
    if prec == rr.FP64:
        e = 0.5 * np.real(np.sum(p.T * (h + f)))
        er = 0.5 * np.real(np.sum(p.T * (h + fr)))
        assert abs(e - er) < 1e-10
        assert np.max(np.abs(f - f.conj().T)) < 1e-12 * np.max(np.abs(f))


def test_x2c_fock_ri_kramers_symmetry(ctx):
    # For a Kramers-paired density the Fock matrix is time-reversal symmetric:
    # F_aa = conj(F_bb), F_ab = -conj(F_ba).
    nl, naux = 20, 50
    b, s, x, p, h0 = x2c_inputs(nl, naux, 3, 311)
    h = np.zeros_like(h0)
    ctx.upload(1, b)
    f = ctx.x2c_fock_ri(1, h, x, s, p)
    faa, fab, fba, fbb = f[:nl, :nl], f[:nl, nl:], f[nl:, :nl], f[nl:, nl:]
    scale = np.max(np.abs(f))
    assert np.max(np.abs(faa - fbb.conj())) < 1e-11 * scale
    assert np.max(np.abs(fab + fba.conj())) < 1e-11 * scale


def ukb_inputs(nl, ns, naux, nocc, seed):
    bll = syn.ri_tensor(naux, nl, 0.4, 6.0, seed)
    bss = syn.ri_tensor(naux, ns, 0.2, 6.0, seed + 1)
    n = 2 * nl + 2 * ns
    g = syn.rng(seed + 2)
    c = g.standard_normal((n, nocc)) + 1j * g.standard_normal((n, nocc))
    occ = 0.6 + 0.4 * g.random(nocc)
    dens = syn.hermitian(n, seed + 3, 0.3)
    return bll, bss, c, occ, dens


@pytest.mark.parametrize("prec", [rr.FP64, rr.FP32, rr.BF16])
def test_ri_fock_two_electron_direct(ctx, oracle, prec):
    nl, ns, naux, nocc = 17, 29, 40, 11
    bll, bss, c, occ, dens = ukb_inputs(nl, ns, naux, nocc, 401)
    ctx.upload(2, bll, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    ctx.upload(3, bss, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    g = ctx.ri_fock_two_electron_direct(2, 3, nl, ns, dens, c, occ, precision=prec)
    gr = oracle.ri_fock_two_electron_direct(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), dens, c, occ)
    assert rel(g, gr) < TOL[prec]


def test_ri_fock_two_electron_direct_shape_errors(ctx):
    bll = syn.ri_tensor(10, 4, 0.3, 3.0, 1)
    bss = syn.ri_tensor(11, 5, 0.3, 3.0, 2)
    ctx.upload(4, bll)
    ctx.upload(5, bss)
    n = 18
    with pytest.raises(rr.GpuError, match="do not share the same aux dimension"):
        ctx.ri_fock_two_electron_direct(4, 5, 4, 5, np.zeros((n, n), complex), np.zeros((n, 1), complex), [1.0])


def test_rkb_fock_ri_embedding(ctx, oracle):
    nl, ns, naux, nocc = 9, 14, 30, 6
    bll, bss, _, _, _ = ukb_inputs(nl, ns, naux, nocc, 501)
    n_ukb = 2 * nl + 2 * ns
    n_rkb = 4 * nl
    g = syn.rng(502)
    v = (g.standard_normal((n_ukb, n_rkb)) + 1j * g.standard_normal((n_ukb, n_rkb))) * 0.3
    sr = v.conj().T @ v + 0.5 * np.eye(n_rkb)
    w, u = np.linalg.eigh(sr)
    x = u / np.sqrt(w)[None, :]
    q = syn.unitary(n_rkb, 503)[:, :nocc]
    c = x @ q
    p = c @ c.conj().T
    h = syn.hermitian(n_rkb, 504)
    ctx.upload(6, bll)
    ctx.upload(7, bss)
    f = ctx.rkb_fock_ri(6, 7, nl, ns, h, v, sr, x, p)
    fr = oracle.rkb_fock_ri(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), v, sr, x, h, p)
    assert rel(f, fr) < 1e-11


@pytest.mark.parametrize("zero_ll,zero_ss,prec", [(20, None, rr.FP64), (None, 23, rr.FP64), (1, 35, rr.FP64),
                                                   (20, None, rr.BF16), (None, 9, rr.FP32)])
def test_rkb_shared_aux_with_zero_trailing_rows(ctx, oracle, zero_ll, zero_ss, prec):
    # unequal aux sets as one shared index: a tensor is zero beyond its own aux count
    nl, ns, naux, nocc = 8, 12, 36, 5
    bll, bss, c, occ, dens = ukb_inputs(nl, ns, naux, nocc, 601)
    if zero_ll is not None:
        bll[zero_ll:] = 0.0
    if zero_ss is not None:
        bss[zero_ss:] = 0.0
    cp = rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16
    ctx.upload(8, bll, cp)
    ctx.upload(9, bss, cp)
    g = ctx.ri_fock_two_electron_direct(8, 9, nl, ns, dens, c, occ, precision=prec)
    gr = oracle.ri_fock_two_electron_direct(nl, ns, bll.reshape(naux, -1), bss.reshape(naux, -1), dens, c, occ)
    assert rel(g, gr) < {rr.FP64: 1e-11, rr.FP32: 1e-5, rr.BF16: 3e-2}[prec]
    if prec == rr.FP64:
        # the same tensors with every row live (trailing rows set to a tiny value) agree
        bll2, bss2 = bll.copy(), bss.copy()
        for z, b2 in ((zero_ll, bll2), (zero_ss, bss2)):
            if z is not None:
                b2[-1, 0, 0] = 1e-300
        ctx.upload(8, bll2)
        ctx.upload(9, bss2)
        g2 = ctx.ri_fock_two_electron_direct(8, 9, nl, ns, dens, c, occ)
        assert rel(g2, g) < 1e-13


def test_x2c_with_zero_trailing_aux_rows(ctx, oracle):
    nl, naux = 30, 80
    b, s, x, p, h = x2c_inputs(nl, naux, 4, 611)
    b[50:] = 0.0
    ctx.upload(10, b)
    f = ctx.x2c_fock_ri(10, h, x, s, p)
    fr = oracle.x2c_fock_ri(b.reshape(naux, -1), x, s, h, p)
    assert rel(f, fr) < 1e-11
