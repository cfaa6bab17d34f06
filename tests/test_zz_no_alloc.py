"""Allocation-free J/K build (rr_ri_jk_factor_noalloc) against rr_ri_jk_factor through the C ABI.

Two calls in a row on the same device, the second on a smaller tensor and factor, each compared
with the allocating entry point on the same inputs (the same kernels in the same order: equal
bitwise); a shape beyond the static scratch region is rejected with an error.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _case(ctx, slot, naux, n, r, seed, prec):
    import rerdmft_amd as rr
    ctx.generate(slot, naux, n, 1.0, 20.0, seed, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    l = torch.randn((n, r), dtype=torch.float64, device="cuda:0", generator=g)
    ja, ka = (torch.empty((n, n), dtype=torch.float64, device="cuda:0") for _ in range(2))
    jb, kb = (torch.empty((n, n), dtype=torch.float64, device="cuda:0") for _ in range(2))
    ctx.ri_jk_factor(slot, None, l, precision=prec, j_out=ja, k_out=ka)
    ctx.ri_jk_factor_noalloc(slot, l, precision=prec, j_out=jb, k_out=kb)
    torch.cuda.synchronize()
    return ja, ka, jb, kb


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_noalloc_equals_allocating_build_twice(ctx, prec):
    print(f"noalloc: precision {prec}, n_aux=800 n=200 r=40 then n_aux=300 n=24 r=5", flush=True)
    for slot, (naux, n, r) in ((8, (800, 200, 40)), (9, (300, 24, 5))):
        ja, ka, jb, kb = _case(ctx, slot, naux, n, r, 20261211 + n, prec)
        assert torch.equal(ja, jb) and torch.equal(ka, kb)
        assert float(ka.abs().max()) > 0.0 and float(ja.abs().max()) > 0.0
        ctx.release(slot)


def test_noalloc_rejects_a_shape_beyond_the_scratch(ctx):
    import rerdmft_amd as rr
    ctx.generate(8, 4000, 500, 1.0, 20.0, 7, rr.COPY_FP64)
    l = torch.zeros((500, 64), dtype=torch.float64, device="cuda:0")
    j, k = (torch.empty((500, 500), dtype=torch.float64, device="cuda:0") for _ in range(2))
    with pytest.raises(rr.GpuError, match="scratch region"):
        ctx.ri_jk_factor_noalloc(8, l, j_out=j, k_out=k)
    ctx.release(8)
