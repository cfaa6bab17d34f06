"""ctypes front-end of the CPU oracle (oracle/rr_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg of
bench.py, always as the checker. The GPU library (rerdmft_amd) never imports this module.

Arrays are numpy, row-major, float64 or complex128 (complex passed as interleaved (re, im)).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_sz = ctypes.c_size_t
_dp = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


def has_mkl():
    return bool(lib().oracle_has_mkl())


def max_threads():
    return int(lib().oracle_max_threads())


def nonrel_fock_ri(eri3, x_large, overlap, h_core, density, return_jk=False):
    """nonRelFockMatrix(h, RiNonRelEriSource, P) -- src/NON_REL/NonRelFockMatrixRi.cpp:27-100."""
    eri3, x, s, h, d = map(_f64, (eri3, x_large, overlap, h_core, density))
    naux, n2 = eri3.shape
    n = h.shape[0]
    assert n * n == n2
    f = np.empty((n, n))
    j = np.empty((n, n))
    k = np.empty((n, n))
    _check(lib().oracle_nonrel_fock_ri(_sz(n), _sz(naux), _p(eri3), _p(x), _sz(x.shape[1]), _p(s), _p(h), _p(d),
                                       _p(f), _p(j), _p(k)))
    return (f, j, k) if return_jk else f


def ri_jk_factor(eri3, d, lfac, rfac=None):
    """J = sum_P tr(B_P D) B_P and K = sum_P B_P^T (L R^T) B_P (closed form, no recovery)."""
    eri3, d, lfac = map(_f64, (eri3, d, lfac))
    rfac = lfac if rfac is None else _f64(rfac)
    naux, n2 = eri3.shape
    n = d.shape[0]
    j = np.empty((n, n))
    k = np.empty((n, n))
    _check(lib().oracle_ri_jk_factor(_sz(n), _sz(naux), _p(eri3), _p(d), _p(lfac), _p(rfac), _sz(lfac.shape[1]),
                                     _p(j), _p(k)))
    return j, k


def x2c_fock_ri(eri3, x_large, overlap, h, density):
    """x2cFockMatrix(h, RiX2cEriSource, P) -- src/X2C_DHF/X2C_FockMatrixRi.cpp:35-162."""
    eri3 = _f64(eri3)
    x, s, h, d = map(_c128, (x_large, overlap, h, density))
    naux = eri3.shape[0]
    nl = h.shape[0] // 2
    f = np.empty_like(h)
    _check(lib().oracle_x2c_fock_ri(_sz(nl), _sz(naux), _p(eri3), _p(x), _sz(x.shape[1]), _p(s), _p(h), _p(d), _p(f)))
    return f


def ri_fock_two_electron_direct(nl, ns, eri3_ll, eri3_ss, density_ukb, c_occ_ukb, occupations):
    """riFockTwoElectronDirect -- src/UKB/UkbFockMatrixRi.cpp:35-183."""
    ell, ess = _f64(eri3_ll), _f64(eri3_ss)
    d, c = _c128(density_ukb), _c128(c_occ_ukb)
    occ = _f64(occupations)
    n = 2 * nl + 2 * ns
    g = np.empty((n, n), dtype=np.complex128)
    _check(lib().oracle_ri_fock_two_electron_direct(_sz(nl), _sz(ns), _sz(ell.shape[0]), _p(ell), _p(ess), _p(d), _p(c),
                                                    _sz(occ.shape[0]), _p(occ), _p(g)))
    return g


def rkb_fock_ri(nl, ns, eri3_ll, eri3_ss, v_total, s_full, x_full, h_rkb, density):
    """rkbFockMatrix(h, RiDirectEriSource, P) -- src/UKB/UkbFockMatrixRi.cpp:185-216."""
    ell, ess = _f64(eri3_ll), _f64(eri3_ss)
    v, s, x, h, d = map(_c128, (v_total, s_full, x_full, h_rkb, density))
    n_rkb = h.shape[0]
    f = np.empty_like(h)
    _check(lib().oracle_rkb_fock_ri(_sz(nl), _sz(ns), _sz(ell.shape[0]), _p(ell), _p(ess), _p(v), _sz(n_rkb), _p(s),
                                    _p(x), _sz(x.shape[1]), _p(h), _p(d), _p(f)))
    return f


def build_ri_mo_eri(nl, ns, eri3_ll, eri3_ss, c_ukb):
    """buildRiMoEri -- src/UKB/RiMoEri.cpp:64-86. Returns (naux, n_total, n_total) complex."""
    ell, ess, c = _f64(eri3_ll), _f64(eri3_ss), _c128(c_ukb)
    naux = ell.shape[0]
    nt = c.shape[1]
    out = np.empty((naux, nt, nt), dtype=np.complex128)
    _check(lib().oracle_build_ri_mo_eri(_sz(nl), _sz(ns), _sz(naux), _p(ell), _p(ess), _p(c), _sz(nt), _p(out)))
    return out


def ri_mo_eri_rotated(b, u):
    """RiMoEri::rotated -- src/UKB/RiMoEri.cpp:88-110."""
    b, u = _c128(b), _c128(u)
    naux, nt = b.shape[0], u.shape[0]
    out = np.empty((naux, nt, nt), dtype=np.complex128)
    _check(lib().oracle_ri_mo_eri_rotated(_sz(nt), _sz(naux), _p(b), _p(u), _p(out)))
    return out


def build_ri_nonrel_spin_mo_eri(eri3, c_spatial):
    """buildRiNonRelSpinMoEri -- src/NON_REL/NonRelSpinRiMoEri.cpp:13-43. (naux, n, n) real."""
    eri3, c = _f64(eri3), _f64(c_spatial)
    naux, n = eri3.shape[0], c.shape[0]
    out = np.empty((naux, n, n))
    _check(lib().oracle_build_ri_nonrel_spin_mo_eri(_sz(n), _sz(naux), _p(eri3), _p(c), _p(out)))
    return out


def ri_nonrel_spin_mo_eri_rotated(b, u):
    """RiNonRelSpinMoEri::rotated -- src/NON_REL/NonRelSpinRiMoEri.cpp:45-95."""
    b, u = _f64(b), _f64(u)
    naux, n = b.shape[0], u.shape[0] // 2
    out = np.empty((naux, n, n))
    _check(lib().oracle_ri_nonrel_spin_mo_eri_rotated(_sz(n), _sz(naux), _p(b), _p(u), _p(out)))
    return out


def nonrel_fock_cholesky(vectors, h, p):
    """nonRelFockMatrix(h, AoCholesky, P) -- src/Utils/AoCholesky.cpp:108-149."""
    v, h, p = _f64(vectors), _f64(h), _f64(p)
    n = h.shape[0]
    f = np.empty((n, n))
    _check(lib().oracle_nonrel_fock_cholesky(_sz(n), _sz(v.shape[0]), _p(v), _p(h), _p(p), _p(f)))
    return f


def x2c_fock_cholesky(vectors, h, p):
    """x2cFockMatrix(h, AoCholesky, P) -- src/Utils/AoCholesky.cpp:151-217."""
    v = _f64(vectors)
    h, p = _c128(h), _c128(p)
    n = h.shape[0] // 2
    f = np.empty_like(h)
    _check(lib().oracle_x2c_fock_cholesky(_sz(n), _sz(v.shape[0]), _p(v), _p(h), _p(p), _p(f)))
    return f


def cholesky_to_mo_spatial(vectors, c):
    """Spatial block of aoCholeskyToMoSpinOrbital -- src/Utils/AoCholesky.cpp:219-243."""
    v, c = _f64(vectors), _f64(c)
    nchol, n, nmo = v.shape[0], c.shape[0], c.shape[1]
    out = np.empty((nchol, nmo, nmo))
    _check(lib().oracle_cholesky_to_mo_spatial(_sz(n), _sz(nchol), _p(v), _p(c), _sz(nmo), _p(out)))
    return out


def cholesky_to_mo_spinor(vectors, c):
    """aoCholeskyToMoSpinor -- src/Utils/AoCholesky.cpp:270-295."""
    v, c = _f64(vectors), _c128(c)
    nchol, nmo = v.shape[0], c.shape[1]
    n = c.shape[0] // 2
    out = np.empty((nchol, nmo, nmo), dtype=np.complex128)
    _check(lib().oracle_cholesky_to_mo_spinor(_sz(n), _sz(nchol), _p(v), _p(c), _sz(nmo), _p(out)))
    return out


def eig_hermitian(a):
    a = _c128(a)
    n = a.shape[0]
    w = np.empty(n)
    v = np.empty((n, n), dtype=np.complex128)
    _check(lib().oracle_eig_hermitian(_sz(n), _p(a), _p(w), _p(v)))
    return w, v


# ---- functional stage (rr_oracle_sigma.cpp) --------------------------------------------------------
def _cx(a, cplx):
    return _c128(a) if cplx else _f64(a)


def _check_sigma(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_sigma_last_error().decode())


def ri_hx_fock_gemm(b, h, occ, h2, x2, pair_of=None, l1=None, l2=None):
    """riHxFockGemm (src/Hessian_opt/RiGemmFock.cpp:53-115)."""
    cplx = np.iscomplexobj(b) or np.iscomplexobj(h)
    n, naux = h.shape[0], b.shape[0]
    bb, hh = _cx(b, cplx), _cx(h, cplx)
    oc, a2, b2 = _f64(occ), _f64(h2), _f64(x2)
    pa = None if pair_of is None else np.ascontiguousarray(np.asarray(pair_of, dtype=np.int64))
    l1a, l2a = (None, None) if pair_of is None else (_f64(l1), _f64(l2))
    f = np.zeros((n, n), dtype=np.complex128 if cplx else np.float64)
    lib().oracle_sigma_last_error.restype = ctypes.c_char_p
    _check_sigma(lib().oracle_ri_hx_fock_gemm(ctypes.c_int(int(cplx)), _sz(n), _sz(naux), _p(bb), _p(hh), _p(oc),
                                              _p(a2), _p(b2), _p(pa), _p(l1a), _p(l2a), _p(f)))
    return f


def ri_sigma_toolkit(b, kappa):
    """buildRiSigmaHessianToolkit (src/Hessian_opt/RiSigmaHessianToolkit.cpp:25-56)."""
    cplx = np.iscomplexobj(b)
    naux, n = b.shape[0], kappa.shape[0]
    dt = np.complex128 if cplx else np.float64
    bb, kk = _cx(b, cplx), _f64(kappa)
    kb1, kb2 = np.zeros((naux, n, n), dt), np.zeros((naux, n, n), dt)
    bd, k1, k2 = np.zeros((naux, n), dt), np.zeros((naux, n), dt), np.zeros((naux, n), dt)
    lib().oracle_ri_sigma_toolkit(ctypes.c_int(int(cplx)), _sz(n), _sz(naux), _p(bb), _p(kk), _p(kb1), _p(kb2),
                                  _p(bd), _p(k1), _p(k2))
    return kb1, kb2, bd, k1, k2


def ri_pair_contract(target, sign, n, naux, f, g):
    """pairContract (src/Hessian_opt/RiSigmaGemm.cpp:76-101); f, g dicts (s, idx, w, row, post)."""
    cplx = np.iscomplexobj(target)
    keep = []

    def fields(d):
        out = []
        for key in ("s", "idx", "w", "row", "post"):
            v = d.get(key)
            if v is None:
                out.append(ctypes.c_void_p(0))
                continue
            dt = np.uint32 if key == "idx" else ((np.complex128 if cplx else np.float64) if key in ("s", "row")
                                                 else np.float64)
            a = np.ascontiguousarray(v, dtype=dt)
            keep.append(a)
            out.append(_p(a))
        return out

    t = _cx(target, cplx)
    lib().oracle_ri_pair_contract(ctypes.c_int(int(cplx)), _sz(n), _sz(naux), ctypes.c_double(sign), *fields(f),
                                  *fields(g), _p(t))
    return t


def ri_diag_contract(target, s, ua=None, ub=None):
    """diagContract (src/Hessian_opt/RiSigmaGemm.cpp:103-121)."""
    cplx = np.iscomplexobj(target)
    n, naux = target.shape[0], s.shape[0]
    t = _cx(target, cplx)
    ss = _cx(s, cplx)
    a = None if ua is None else _cx(ua, cplx)
    b_ = None if ub is None else _cx(ub, cplx)
    lib().oracle_ri_diag_contract(ctypes.c_int(int(cplx)), _sz(n), _sz(naux), _p(ss), _p(a), _p(b_), _p(t))
    return t


def ri_accum_u(u, alpha, d, wm):
    """accumU (src/Hessian_opt/RiSigmaGemm.cpp:123-132)."""
    cplx = np.iscomplexobj(u)
    naux, n = u.shape
    uu, dd, ww = _cx(u, cplx), _cx(d, cplx), _f64(wm)
    lib().oracle_ri_accum_u(ctypes.c_int(int(cplx)), _sz(n), _sz(naux), ctypes.c_double(alpha), _p(dd), _p(ww), _p(uu))
    return uu


def sigma_hessian_vector_ri(h, b, occ, hc, xc, pairs, v, fock=None, pair_of=None, l1=None, l2=None, hx=False):
    """jkOnlySigma(Joint)HessianVectorRi (JkOnlySigmaHessianVectorRi.cpp:504-573) and
    hartreeExchangeSigma(Joint)HessianVectorRi (HartreeExchangeSigmaHessianVectorRi.cpp:442-539).
    Complex h/b select the joint [t; y] form."""
    cplx = np.iscomplexobj(b) or np.iscomplexobj(h)
    n, naux = h.shape[0], b.shape[0]
    pp = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1, 2))
    npairs = pp.shape[0]
    hh, bb = _cx(h, cplx), _cx(b, cplx)
    oc, a, x = _f64(occ), _f64(hc), _f64(xc)
    ff = None if fock is None else _cx(fock, cplx)
    pa = None if pair_of is None else np.ascontiguousarray(np.asarray(pair_of, dtype=np.int64))
    l1a, l2a = (None, None) if pair_of is None else (_f64(l1), _f64(l2))
    vv = _f64(v)
    w = np.zeros(npairs * (2 if cplx else 1))
    lib().oracle_sigma_last_error.restype = ctypes.c_char_p
    _check_sigma(lib().oracle_sigma_hessian_vector_ri(ctypes.c_int(int(cplx)), ctypes.c_int(int(hx)), _sz(n), _p(hh),
                                                      _p(bb), _sz(naux), _p(oc), _p(a), _p(x), _p(ff), _p(pa),
                                                      _p(l1a), _p(l2a), _sz(npairs), _p(pp), _p(vv), _p(w)))
    return w


# ---- one-time RI tensor preparation (rr_oracle.cpp) ------------------------------------------------
def ri_three_center_tensor(raw, x):
    """riThreeCenterTensor's contraction (src/AO_ints/ThreeCenterIntegrals.cpp:219-236)."""
    naux_raw = raw.shape[0]
    n2 = int(np.prod(raw.shape[1:]))
    n = int(round(np.sqrt(n2)))
    kept = x.shape[1]
    r, xx = _f64(raw), _f64(x)
    out = np.zeros((kept, n2))
    _check(lib().oracle_ri_three_center_tensor(_sz(n), _sz(naux_raw), _p(r), _p(xx), _sz(kept), _p(out)))
    return out


def transform_ao_legs(eri3, n_ao, t):
    """transformRiThreeCenterAoLegs (src/AO_ints/ThreeCenterIntegrals.cpp:191-217)."""
    naux, n_out = eri3.shape[0], t.shape[1]
    e, tt = _f64(eri3), _f64(t)
    out = np.zeros((naux, n_out * n_out))
    _check(lib().oracle_transform_ao_legs(_sz(naux), _sz(n_ao), _p(e), _p(tt), _sz(n_out), _p(out)))
    return out


def canonical_orthogonalize(s, threshold=1e-10):
    """canonicalOrthogonalize (src/Linear_Algebra/LinearAlgebra.cpp:267-310)."""
    n = s.shape[0]
    ss = _f64(s)
    x = np.zeros(n * n)
    kept = _sz()
    _check(lib().oracle_canonical_orthogonalize(_sz(n), _p(ss), ctypes.c_double(threshold), _p(x), ctypes.byref(kept)))
    return x[: n * kept.value].reshape(n, kept.value)


# ---- 4c RKB Cholesky (rr_oracle.cpp) ---------------------------------------------------------------
def _rkb_args(large, small0, small1, cross_f, cross_b):
    arrs = [_f64(large), _c128(small0), _c128(small1), _c128(cross_f), _c128(cross_b)]
    return arrs, [_p(a) for a in arrs]


def rkb_fock_cholesky(large, small0, small1, cross_f, cross_b, h, p):
    """rkbFockMatrix(h_rkb, RkbCholesky, P) -- src/C4_DHF/RkbCholesky.cpp:352-448."""
    nchol = large.shape[0]
    nl = int(round(np.sqrt(np.prod(large.shape[1:]))))
    keep, ptrs = _rkb_args(large, small0, small1, cross_f, cross_b)
    hh, pp = _c128(h), _c128(p)
    f = np.zeros((4 * nl, 4 * nl), dtype=np.complex128)
    _check(lib().oracle_rkb_fock_cholesky(_sz(nl), _sz(nchol), *ptrs, _p(hh), _p(pp), _p(f)))
    return f


def rkb_cholesky_to_mo(large, small0, small1, cross_f, cross_b, c, n_negative=0, threshold=1e-10):
    """rkbCholeskyToMo -- src/C4_DHF/RkbCholesky.cpp:450-523 (recompression :244-283)."""
    nchol = large.shape[0]
    nl = int(round(np.sqrt(np.prod(large.shape[1:]))))
    keep, ptrs = _rkb_args(large, small0, small1, cross_f, cross_b)
    cc = _c128(c)
    nmo = cc.shape[1]
    m = nmo - n_negative
    per = m * m if n_negative == 0 else nmo * nmo
    out = np.zeros((nchol, per), dtype=np.complex128)
    nout = _sz()
    _check(lib().oracle_rkb_cholesky_to_mo(_sz(nl), _sz(nchol), *ptrs, _p(cc), _sz(nmo), _sz(n_negative),
                                           ctypes.c_double(threshold), _p(out), _sz(nchol), ctypes.byref(nout)))
    return out[: nout.value]
