"""Every benchmark configuration at its stated shape on the GPU.

Each configuration runs once at full size, checked through properties that hold at any size (K and
F symmetric or Hermitian, time-reversal symmetry for Kramers-paired densities, linearity in the
density, agreement of independent code paths, an FP64 incremental sequence against the full build),
and once on a contiguous block of the same tensor's aux rows against the oracle: J and K are sums
over the aux index, so a block is an exact J/K problem of its own with every other dimension at the
configuration's size. The device tensors come from Context.generate; the block's rows are
reproduced on the host by synthetic.device_tensor_rows (the same counter-based generator).

Tolerances: FP64 1e-10 relative to the largest element (the oracle and the device sum in different
orders over up to 10^6 terms); properties 1e-12; FP32 / BF16 incremental builds 1e-5 / 3e-2 of the
FP64 incremental build.
"""
import numpy as np
import pytest

import rerdmft_amd as rr
from rerdmft_amd import synthetic as syn

pytestmark = pytest.mark.gpu
ALL = rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1e-300, np.max(np.abs(np.asarray(b)))))


def note(msg):
    print(msg, flush=True)


def test_config1_kernel_level_water(ctx, oracle):
    """B[300][24][24], N(0, 0.1) exp(-|mu-nu|/8), 5 doubly occupied orbitals, S = X = I."""
    note("config 1: n=24 n_aux=300 n_occ=5")
    b = syn.ri_tensor(300, 24, 0.1, 8.0, syn.BASE_SEED)
    c = syn.orthonormal_occ(24, 5, syn.BASE_SEED + 100)
    p = syn.density(c)
    h = syn.symmetric(24, syn.BASE_SEED + 200)
    eye = np.eye(24)
    ctx.upload(0, b, ALL)
    f = ctx.nonrel_fock_ri(0, h, eye, eye, p)
    fr = oracle.nonrel_fock_ri(b.reshape(300, -1), eye, eye, h, p)
    assert rel(f, fr) < 1e-12
    j, k = ctx.ri_jk_factor(0, None, np.sqrt(2.0) * c)
    jr, kr = oracle.ri_jk_factor(b.reshape(300, -1), p, np.sqrt(2.0) * c)
    assert rel(j, jr) < 1e-12 and rel(k, kr) < 1e-12
    ctx.release(0)


def test_config2_large_real(ctx, oracle):
    """B[4000][1000][1000] sym N(0,1) exp(-|mu-nu|/50), n_occ = 200; incremental builds."""
    import torch
    n, naux, no, seed, blk = 1000, 4000, 200, syn.BASE_SEED + 1, 160
    dev = torch.device("cuda:0")
    note(f"config 2: block of {blk} aux rows vs the oracle")
    c = syn.orthonormal_occ(n, no, seed + 7)
    l = np.sqrt(2.0) * c
    ctx.generate(0, blk, n, 1.0, 50.0, seed, rr.COPY_FP64)
    b = syn.device_tensor_rows(n, 1.0, 50.0, seed, 0, blk).reshape(blk, -1)
    j, k = ctx.ri_jk_factor(0, None, l)
    jr, kr = oracle.ri_jk_factor(b, l @ l.T, l)
    del b
    assert rel(j, jr) < 1e-10 and rel(k, kr) < 1e-10
    jd, _ = ctx.ri_jk_factor(0, l @ l.T, l)  # Coulomb from the density (two passes over B)
    assert rel(jd, jr) < 1e-10
    ctx.release(0)

    note("config 2: full size, properties")
    ctx.generate(1, naux, n, 1.0, 50.0, seed, ALL)
    ld = torch.tensor(l, device=dev)
    J, K = (torch.empty((n, n), dtype=torch.float64, device=dev) for _ in range(2))
    ctx.ri_jk_factor(1, None, ld, j_out=J, k_out=K)
    jf, kf = J.cpu().numpy(), K.cpu().numpy()
    assert rel(kf, kf.T) < 1e-12 and rel(jf, jf.T) < 1e-12
    # linearity in D = L L^T over a split of the occupied columns; J by two passes over B
    j1, k1 = ctx.ri_jk_factor(1, None, ld[:, :120].contiguous())
    j2, k2 = ctx.ri_jk_factor(1, None, ld[:, 120:].contiguous())
    assert rel(j1 + j2, jf) < 1e-12 and rel(k1 + k2, kf) < 1e-12
    jd, _ = ctx.ri_jk_factor(1, torch.tensor(l @ l.T, device=dev), None, want_k=False)
    assert rel(jd, jf) < 1e-12
    # incremental builds along an orbital path: FP64 increments sum to the full build; FP32 / BF16
    # increments close to the FP64 ones
    g = syn.rng(seed + 8)
    basis, _ = np.linalg.qr(g.standard_normal((n, n)))
    cs = [np.ascontiguousarray(basis[:, :no])]
    for step, angle in enumerate((0.1, 0.01, 0.001)):
        cn = syn.rotate_occ(cs[-1], basis, angle, seed + 20 + step)
        virt, _ = np.linalg.qr(basis[:, no:] - cn @ (cn.T @ basis[:, no:]))
        basis = np.concatenate([cn, virt[:, :n - no]], axis=1)
        cs.append(cn)
    f0j, f0k = ctx.ri_jk_factor(1, None, np.sqrt(2.0) * cs[0])
    accj, acck = f0j.copy(), f0k.copy()
    for a_, b_ in zip(cs[1:], cs[:-1]):
        dj, dk = ctx.ri_jk_delta(1, a_, b_, precision=rr.FP64)
        accj += dj
        acck += dk
        for prec, tol in ((rr.FP32, 1e-5), (rr.BF16, 3e-2)):
            djl, dkl = ctx.ri_jk_delta(1, a_, b_, precision=prec)
            assert rel(djl, dj) < tol and rel(dkl, dk) < tol
    fj, fk = ctx.ri_jk_factor(1, None, np.sqrt(2.0) * cs[-1])
    assert rel(accj, fj) < 1e-11 and rel(acck, fk) < 1e-11
    ctx.release(1)


def test_config3_x2c_and_mo_transform(ctx, oracle):
    """X2C: n = 400 spatial (800 spinors), n_aux = 2400, 80 Kramers pairs; AO->MO to 800 x 800."""
    nl, naux, npairs, seed, blk = 400, 2400, 80, syn.BASE_SEED + 2, 48
    note(f"config 3: block of {blk} aux rows vs the oracle")
    q = syn.kramers_pair_occ(nl, npairs, seed + 10)
    p = q @ q.conj().T
    eye = np.eye(2 * nl, dtype=np.complex128)
    h = syn.hermitian(2 * nl, seed + 11)
    ctx.generate(2, blk, nl, 1.0, 40.0, seed, rr.COPY_FP64)
    b = syn.device_tensor_rows(nl, 1.0, 40.0, seed, 0, blk).reshape(blk, -1)
    f = ctx.x2c_fock_ri(2, h, eye, eye, p)
    fr = oracle.x2c_fock_ri(b, eye, eye, h, p)
    assert rel(f, fr) < 1e-10
    u = syn.unitary(2 * nl, seed + 12)
    mo = ctx.build_ri_mo_eri(2, 2, nl, 0, u)
    mor = oracle.build_ri_mo_eri(nl, 0, b, np.zeros((blk, 0)), u)
    assert rel(mo.reshape(blk, -1), mor.reshape(blk, -1)) < 1e-10
    del b, mo, mor
    ctx.release(2)

    note("config 3: full size, properties")
    ctx.generate(3, naux, nl, 1.0, 40.0, seed, ALL)
    hz = np.zeros_like(h)
    for prec in (rr.FP64, rr.FP32):
        f = ctx.x2c_fock_ri(3, hz, eye, eye, p, precision=prec)
        s = np.max(np.abs(f))
        assert np.max(np.abs(f - f.conj().T)) < 1e-12 * s
        faa, fab, fba, fbb = f[:nl, :nl], f[:nl, nl:], f[nl:, :nl], f[nl:, nl:]
        tol = 1e-11 if prec == rr.FP64 else 1e-5
        assert np.max(np.abs(faa - fbb.conj())) < tol * s and np.max(np.abs(fab + fba.conj())) < tol * s
    ctx.release(3)


def test_config4_rkb_shared_aux(ctx, oracle):
    """4c RKB: nL = 300, nS = 700 (n_ukb = 2000), shared n_aux = 4200 with the LL tensor zero past row
    1800, 100 Kramers pairs."""
    import torch
    nl, ns, naux, nll, npairs, seed = 300, 700, 4200, 1800, 100, syn.BASE_SEED + 3
    n = 2 * nl + 2 * ns
    c = syn.kramers_pair_ukb(nl, ns, npairs, seed + 10)
    p = c @ c.conj().T
    occ = np.ones(2 * npairs)
    p0, p1 = 1780, 1840  # a block across the end of the LL tensor's live rows
    note(f"config 4: aux rows [{p0}, {p1}) vs the oracle")
    bll_full = syn.device_tensor_rows(nl, 1.0, 30.0, seed, p0, p1)
    bll_full[nll - p0:] = 0.0
    bss = syn.device_tensor_rows(ns, 1.0, 70.0, seed + 1, p0, p1)
    ctx.upload(4, bll_full, rr.COPY_FP64)
    ctx.upload(5, bss, rr.COPY_FP64)
    g = ctx.ri_fock_two_electron_direct(4, 5, nl, ns, p, c, occ)
    gr = oracle.ri_fock_two_electron_direct(nl, ns, bll_full.reshape(p1 - p0, -1), bss.reshape(p1 - p0, -1), p, c, occ)
    assert rel(g, gr) < 1e-10
    ctx.release(4)
    ctx.release(5)
    del bll_full, bss

    note("config 4: full size, properties")
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    idx = torch.arange(nl, device=dev, dtype=torch.float64)
    bl = torch.zeros((naux, nl, nl), dtype=torch.float64, device=dev)
    z = torch.randn((nll, nl, nl), dtype=torch.float64, device=dev, generator=gen)
    bl[:nll] = 0.5 * (z + z.transpose(1, 2)) * torch.exp(-(idx[:, None] - idx[None, :]).abs() / 30.0)
    del z
    ctx.upload(6, bl, ALL)
    del bl
    ctx.generate(7, naux, ns, 1.0, 70.0, seed + 1, ALL)
    u = syn.time_reversal_ukb(nl, ns)
    for prec in (rr.FP64, rr.FP32, rr.BF16):
        g = ctx.ri_fock_two_electron_direct(6, 7, nl, ns, p, c, occ, precision=prec)
        s = np.max(np.abs(g))
        tol = {rr.FP64: 1e-12, rr.FP32: 1e-5, rr.BF16: 1e-2}[prec]
        assert np.max(np.abs(g - g.conj().T)) < max(tol, 1e-12) * s
        assert np.max(np.abs(g - u @ g.conj() @ u.T)) < tol * s
    ctx.release(6)
    ctx.release(7)


def test_config5a_shared_tensor_densities(ctx, oracle):
    """64 trial densities D_k = C_occ T_k C_virt^T (n_occ = 100, n_virt = 700, n = 800, n_aux = 3200)."""
    import torch
    n, no, naux, nd, seed = 800, 100, 3200, 64, syn.BASE_SEED + 4
    dev = torch.device("cuda:0")
    g = syn.rng(seed)
    q, _ = np.linalg.qr(g.standard_normal((n, n)))
    c_occ, c_virt = np.ascontiguousarray(q[:, :no]), q[:, no:]
    t = g.standard_normal((nd, no, n - no))
    rf = np.ascontiguousarray(np.einsum("va,kia->kvi", c_virt, t))  # R_k = C_virt T_k^T, (nd, n, no)
    note("config 5a: all 64 densities at full size vs single builds")
    ctx.generate(8, naux, n, 1.0, 40.0, seed, rr.COPY_FP64)
    lo, rfd = torch.tensor(c_occ, device=dev), torch.tensor(rf, device=dev)
    js, ks = (torch.empty((nd, n, n), dtype=torch.float64, device=dev) for _ in range(2))
    ctx.ri_jk_multi(8, lo, rfd, j_out=js, k_out=ks)
    for k in range(nd):
        j1, k1 = ctx.ri_jk_factor(8, None, lo, rfd[k].contiguous())
        assert rel(js[k].cpu().numpy(), j1) < 1e-12 and rel(ks[k].cpu().numpy(), k1) < 1e-12
    ctx.release(8)
    blk = 120
    note(f"config 5a: 3 densities on a block of {blk} aux rows vs the oracle")
    ctx.generate(9, blk, n, 1.0, 40.0, seed, rr.COPY_FP64)
    b = syn.device_tensor_rows(n, 1.0, 40.0, seed, 0, blk).reshape(blk, -1)
    jb, kb, _ = ctx.ri_jk_multi(9, c_occ, rf[:3])
    for k in range(3):
        jr, kr = oracle.ri_jk_factor(b, c_occ @ rf[k].T, c_occ, rf[k])
        assert rel(jb[k], jr) < 1e-10 and rel(kb[k], kr) < 1e-10
    ctx.release(9)


def test_config5b_uneven_batch(ctx, oracle):
    """256 independent builds, n ~ U[20, 200], n_aux = 4n, n_occ = n/5; every molecule vs the oracle."""
    import torch
    g = np.random.default_rng(20261203 + 5)
    ns = [int(x) for x in g.integers(20, 201, 256)]
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261208)
    note(f"config 5b: 256 molecules, n in [{min(ns)}, {max(ns)}]")
    tens = []
    for n in ns:
        x = torch.randn((4 * n, n, n), device=dev, dtype=torch.float64, generator=gen)
        tens.append(0.5 * (x + x.transpose(1, 2)))
    batch = ctx.batch_create(tens, rr.COPY_FP64)
    ls = [torch.randn((n, max(1, n // 5)), device=dev, dtype=torch.float64, generator=gen) for n in ns]
    js, ks, _ = ctx.batch_jk_factor(batch, ls)
    worst = 0.0
    for m, n in enumerate(ns):
        b = tens[m].reshape(4 * n, -1).cpu().numpy()
        l = ls[m].cpu().numpy()
        jr, kr = oracle.ri_jk_factor(b, l @ l.T, l)
        worst = max(worst, rel(js[m], jr), rel(ks[m], kr))
    ctx.batch_destroy(batch)
    note(f"  largest relative deviation over the batch: {worst:.2e}")
    assert worst < 1e-10
