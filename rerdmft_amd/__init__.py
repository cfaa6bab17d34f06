"""rerdmft_amd: MI355X-native RI / Cholesky J/K Fock builds and RI AO->MO transforms.

Python front-end over the C ABI of librerdmft_gpu.so (include/rerdmft_gpu.h). The function names
mirror the reference's C++ overload set (nonRelFockMatrix, x2cFockMatrix, rkbFockMatrix,
riFockTwoElectronDirect, buildRiMoEri, ...), with the same argument meaning and the same shape
errors (raised as RuntimeError). Arrays may be numpy (host) or torch device tensors; results are
returned as numpy arrays unless an `out=` device tensor is given.

There is no CPU fallback: importing succeeds without a GPU, but creating a Context requires the
built library and a visible device, and fails loudly otherwise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librerdmft_gpu.so")

FP64, FP32, BF16 = 0, 1, 2
COPY_FP64, COPY_FP32, COPY_BF16 = 1, 2, 4
PRECISION_NAMES = {FP64: "f64", FP32: "f32", BF16: "bf16"}

# Exported C symbols declared in include/rerdmft_gpu.h (checked by the CPU test suite).
EXPORTED = [
    "rr_create", "rr_destroy", "rr_last_error", "rr_set_tile_config", "rr_set_recovery_method", "rr_timing",
    "rr_tensor_upload", "rr_tensor_generate", "rr_tensor_info", "rr_tensor_release",
    "rr_nonrel_fock_ri", "rr_ri_jk_factor", "rr_ri_jk_factor_noalloc", "rr_ri_jk_multi", "rr_ri_jk_delta", "rr_x2c_fock_ri",
    "rr_ri_fock_two_electron_direct", "rr_rkb_fock_ri", "rr_build_ri_mo_eri", "rr_ri_mo_eri_rotated",
    "rr_ri_mo_eri_positive_block", "rr_build_ri_nonrel_spin_mo_eri", "rr_ri_nonrel_spin_mo_eri_rotated",
    "rr_ri_nonrel_spin_embedded_b", "rr_nonrel_fock_cholesky", "rr_x2c_fock_cholesky",
    "rr_cholesky_to_mo_spatial", "rr_cholesky_to_mo_spinor",
    "rr_ri_hx_fock_gemm", "rr_ri_sigma_toolkit", "rr_ri_pair_contract", "rr_ri_diag_contract", "rr_ri_accum_u",
    "rr_jk_only_sigma_hessian_vector_ri", "rr_jk_only_sigma_joint_hessian_vector_ri",
    "rr_hartree_exchange_sigma_hessian_vector_ri", "rr_hartree_exchange_sigma_joint_hessian_vector_ri",
    "rr_ri_three_center_tensor", "rr_transform_ri_three_center_ao_legs", "rr_canonical_orthogonalize",
    "rr_ri_mo_fock_rkb", "rr_ri_mo_fock_nonrel",
    "rr_rkb_cholesky_upload", "rr_rkb_cholesky_release", "rr_rkb_fock_cholesky", "rr_rkb_fock_cholesky_factor",
    "rr_rkb_cholesky_to_mo",
    "rr_batch_create", "rr_batch_destroy", "rr_batch_jk_factor",
    "rr_mlp_create", "rr_mlp_destroy", "rr_mlp_num_params", "rr_mlp_get_params", "rr_mlp_set_params",
    "rr_mlp_forward", "rr_mlp_loss_grad", "rr_mlp_train_step",
    "rr_scf_commutator_error", "rr_diis_create", "rr_diis_destroy", "rr_diis_clear", "rr_diis_extrapolate",
    "rr_cas_two_electron_operator_ri", "rr_cas_active_integrals_ri", "rr_cas_active_cumulant_fock_ri",
    "rr_cc_physics_block",
]


class PairOp(ctypes.Structure):
    """rr_pair_op (risigma::PairOp): value(P,t,x) = s(P, idx[t*n+x]) * w[t*n+x] * row(P,t), then @ post."""
    _fields_ = [("s", ctypes.c_void_p), ("idx", ctypes.c_void_p), ("w", ctypes.c_void_p), ("row", ctypes.c_void_p),
                ("post", ctypes.c_void_p)]

_lib = None


class GpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load librerdmft_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rerdmft_amd.build`")
        # One HIP runtime per process: PyTorch ships its own libamdhip64 with the same soname, and
        # whichever is loaded first serves both. Load PyTorch's first (when it is installed) so that
        # device tensors created after this library's context see the runtime PyTorch expects.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.rr_last_error.restype = ctypes.c_char_p
        _lib.rr_last_error.argtypes = [ctypes.c_void_p]
        _lib.rr_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        _lib.rr_destroy.argtypes = [ctypes.c_void_p]
    return _lib


_sz = ctypes.c_size_t


def _count(x):
    """Element count of a numpy array or torch tensor (complex values count once)."""
    return int(x.numel()) if type(x).__module__.startswith("torch") else int(np.size(x))


def _need(x, count, what):
    """The C ABI reads or writes exactly `count` elements through a plain pointer; an array of any other
    size would be read or written past its end, so the size is checked here, before the call."""
    if x is not None and _count(x) != count:
        raise ValueError(f"{what}: expected {count} elements, got {_count(x)} (shape {tuple(x.shape)})")


class _Arr:
    """Pointer view of a numpy array or torch tensor (keeps a contiguous copy alive). With `count`,
    the array must hold exactly that many elements."""

    def __init__(self, x, dtype, count=None, what="array"):
        self.torch = False
        if count is not None:
            _need(x, count, what)
        if x is None:
            self.ptr = ctypes.c_void_p(0)
            self.obj = None
            return
        if type(x).__module__.startswith("torch"):
            import torch
            want = {np.float64: torch.float64, np.complex128: torch.complex128, np.float32: torch.float32}[dtype]
            t = x.contiguous()
            if t.dtype != want:
                t = t.to(want)
            self.obj = t
            self.ptr = ctypes.c_void_p(t.data_ptr())
            self.torch = True
        else:
            a = np.ascontiguousarray(x, dtype=dtype)
            self.obj = a
            self.ptr = ctypes.c_void_p(a.ctypes.data)


def _f64(x, count=None, what="array"):
    return _Arr(x, np.float64, count, what)


def _is_complex(*xs):
    for x in xs:
        if type(x).__module__.startswith("torch"):
            if x.is_complex():
                return True
        elif np.iscomplexobj(x):
            return True
    return False


def _c128(x, count=None, what="array"):
    return _Arr(x, np.complex128, count, what)


def _out(out, shape, dtype):
    if out is not None:
        return out, _Arr(out, dtype, int(np.prod(shape)), "out").ptr
    a = np.empty(shape, dtype=dtype)
    return a, ctypes.c_void_p(a.ctypes.data)


class Context:
    """One device context (stream, solver handle, resident 3-index tensors in numbered slots)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().rr_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise GpuError(rc, f"rr_create failed on device {device} (code {rc})")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib().rr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _call(self, name, *args):
        rc = getattr(lib(), name)(self._h, *args)
        if rc != 0:
            raise GpuError(rc, lib().rr_last_error(self._h).decode())

    # ---- configuration / timing ----------------------------------------------------------------
    def set_tile_config(self, cfg):
        self._call("rr_set_tile_config", ctypes.c_int(cfg))

    PHASES = ["gamma", "coulomb", "exchange_half", "exchange_rank_k", "assembly", "recovery"]

    def recovery_method(self, method=-1):
        """Select the occupied-set recovery of the density-taking builds (0 spectral projector with
        eigensolver fallback, 1 eigensolver; -1 keep); returns (projector count, eigensolver count)."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self._call("rr_set_recovery_method", ctypes.c_int(method), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def timing(self, enable=True, reset=False, by_precision=False):
        """Per-phase device time (ms, launches) since the last reset; by_precision splits by operand type."""
        ms = (ctypes.c_double * 24)()
        cnt = (ctypes.c_longlong * 24)()
        self._call("rr_timing", ctypes.c_int(int(enable)), ctypes.c_int(int(reset)), ms, cnt, ctypes.c_int(24))
        if by_precision:
            return {(PRECISION_NAMES[p], k): (ms[8 * p + i], cnt[8 * p + i])
                    for p in range(3) for i, k in enumerate(self.PHASES)}
        return {k: (sum(ms[8 * p + i] for p in range(3)), sum(cnt[8 * p + i] for p in range(3)))
                for i, k in enumerate(self.PHASES)}

    # ---- tensors -------------------------------------------------------------------------------
    def upload(self, slot, eri3, copies=COPY_FP64):
        a = _f64(eri3)
        shape = tuple(eri3.shape)
        naux = shape[0]
        n2 = int(np.prod(shape[1:]))
        n = int(round(np.sqrt(n2)))
        if n * n != n2:
            raise ValueError("eri3 must be (n_aux, n*n) or (n_aux, n, n)")
        self._call("rr_tensor_upload", ctypes.c_int(slot), a.ptr, _sz(naux), _sz(n), ctypes.c_uint(copies))

    def generate(self, slot, naux, n, sigma, decay, seed, copies=COPY_FP64):
        self._call("rr_tensor_generate", ctypes.c_int(slot), _sz(naux), _sz(n), ctypes.c_double(sigma),
                   ctypes.c_double(decay), ctypes.c_uint64(seed), ctypes.c_uint(copies))

    def tensor_info(self, slot):
        na, n, cp = _sz(), _sz(), ctypes.c_uint()
        self._call("rr_tensor_info", ctypes.c_int(slot), ctypes.byref(na), ctypes.byref(n), ctypes.byref(cp))
        return na.value, n.value, cp.value

    def release(self, slot):
        self._call("rr_tensor_release", ctypes.c_int(slot))

    # ---- NON_REL -------------------------------------------------------------------------------
    def nonrel_fock_ri(self, slot, h_core, x_large, overlap, density, precision=FP64, return_jk=False, out=None):
        """nonRelFockMatrix(h, RiNonRelEriSource, P) -- src/NON_REL/NonRelFockMatrixRi.cpp:27-100."""
        n = h_core.shape[0]
        nx = x_large.shape[1]
        h, x = _f64(h_core, n * n, "h_core"), _f64(x_large, n * nx, "x_large")
        s, d = _f64(overlap, n * n, "overlap"), _f64(density, n * n, "density")
        f, fp = _out(out, (n, n), np.float64)
        j = k = None
        jp = kp = ctypes.c_void_p(0)
        if return_jk:
            j, jp = _out(None, (n, n), np.float64)
            k, kp = _out(None, (n, n), np.float64)
        self._call("rr_nonrel_fock_ri", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), h.ptr, x.ptr,
                   _sz(x_large.shape[1]), s.ptr, d.ptr, fp, jp, kp)
        return (f, j, k) if return_jk else f

    def ri_jk_factor(self, slot, density, lfac, rfac=None, precision=FP64, want_j=True, want_k=True, j_out=None,
                     k_out=None):
        """J = sum_P tr(B_P D) B_P, K = sum_P B_P^T (L R^T) B_P."""
        n = lfac.shape[0] if lfac is not None else density.shape[0]
        # density=None: J of D = L R^T from the half-transformed factors (fused Coulomb contraction)
        r = lfac.shape[1] if lfac is not None else 0
        d, l, rr_ = _f64(density, n * n, "density"), _f64(lfac, n * r, "lfac"), _f64(rfac, n * r, "rfac")
        j, jp = _out(j_out, (n, n), np.float64) if want_j else (None, ctypes.c_void_p(0))
        k, kp = _out(k_out, (n, n), np.float64) if want_k else (None, ctypes.c_void_p(0))
        self._call("rr_ri_jk_factor", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), d.ptr, l.ptr, rr_.ptr,
                   _sz(r), jp, kp)
        return j, k

    def ri_jk_factor_noalloc(self, slot, lfac, precision=FP64, j_out=None, k_out=None):
        """Allocation-free J, K of D = L L^T (rr_ri_jk_factor_noalloc): lfac, j_out and k_out are
        device tensors; the workspace is the library's static per-device scratch region."""
        n, r = lfac.shape
        l = _f64(lfac)
        if not (l.torch and j_out is not None and k_out is not None):
            raise ValueError("ri_jk_factor_noalloc: lfac, j_out and k_out must be device tensors")
        j, jp = _out(j_out, (n, n), np.float64)
        k, kp = _out(k_out, (n, n), np.float64)
        self._call("rr_ri_jk_factor_noalloc", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), l.ptr, _sz(r), jp,
                   kp)
        return j, k

    def ri_jk_multi(self, slot, lfac, rfacs, precision=FP64, want_j=True, want_k=True, j_out=None, k_out=None):
        """J_k, K_k of D_k = L R_k^T for a stack of right factors rfacs (n_dens, n, r), sharing L (n, r).
        Returns (J (n_dens, n, n) or None, K or None, replayed)."""
        n, r = lfac.shape
        nd = rfacs.shape[0]
        l, rf = _f64(lfac), _f64(rfacs, nd * n * r, "rfacs")
        j, jp = _out(j_out, (nd, n, n), np.float64) if want_j else (None, ctypes.c_void_p(0))
        k, kp = _out(k_out, (nd, n, n), np.float64) if want_k else (None, ctypes.c_void_p(0))
        rep = ctypes.c_int(0)
        self._call("rr_ri_jk_multi", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), _sz(r), l.ptr, _sz(nd), rf.ptr,
                   jp, kp, ctypes.byref(rep))
        return j, k, bool(rep.value)

    def ri_jk_delta(self, slot, c_new, c_old, precision=FP64, dj_out=None, dk_out=None):
        """dJ, dK of D_new - D_old with D = 2 C C^T (aligned-difference factorization)."""
        n, no = c_new.shape
        a, b = _f64(c_new), _f64(c_old, n * no, "c_old")
        dj, djp = _out(dj_out, (n, n), np.float64)
        dk, dkp = _out(dk_out, (n, n), np.float64)
        self._call("rr_ri_jk_delta", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), _sz(no), a.ptr, b.ptr, djp,
                   dkp)
        return dj, dk

    # ---- X2C -----------------------------------------------------------------------------------
    def x2c_fock_ri(self, slot, h_x2c, x_large, overlap, density, precision=FP64, out=None):
        """x2cFockMatrix(h, RiX2cEriSource, P) -- src/X2C_DHF/X2C_FockMatrixRi.cpp:35-162."""
        n2 = h_x2c.shape[0]
        nx = x_large.shape[1]
        h, x = _c128(h_x2c, n2 * n2, "h_x2c"), _c128(x_large, n2 * nx, "x_large")
        s, d = _c128(overlap, n2 * n2, "overlap"), _c128(density, n2 * n2, "density")
        f, fp = _out(out, (n2, n2), np.complex128)
        self._call("rr_x2c_fock_ri", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n2 // 2), h.ptr, x.ptr,
                   _sz(x_large.shape[1]), s.ptr, d.ptr, fp)
        return f

    # ---- RKB -----------------------------------------------------------------------------------
    def ri_fock_two_electron_direct(self, slot_ll, slot_ss, nl, ns, density_ukb, c_occ_ukb, occupations,
                                    precision=FP64, out=None):
        """riFockTwoElectronDirect -- src/UKB/UkbFockMatrixRi.cpp:35-183."""
        n = 2 * nl + 2 * ns
        no = len(occupations)
        d, c = _c128(density_ukb, n * n, "density_ukb"), _c128(c_occ_ukb, n * no, "c_occ_ukb")
        occ = _f64(occupations)
        g, gp = _out(out, (n, n), np.complex128)
        self._call("rr_ri_fock_two_electron_direct", ctypes.c_int(slot_ll), ctypes.c_int(slot_ss),
                   ctypes.c_int(precision), _sz(nl), _sz(ns), d.ptr, c.ptr, _sz(len(occupations)), occ.ptr, gp)
        return g

    def rkb_fock_ri(self, slot_ll, slot_ss, nl, ns, h_rkb, v_total, s_full, x_full, density, precision=FP64,
                    out=None):
        """rkbFockMatrix(h, RiDirectEriSource, P) -- src/UKB/UkbFockMatrixRi.cpp:185-216."""
        n_rkb = h_rkb.shape[0]
        nu, nx = 2 * nl + 2 * ns, x_full.shape[1]
        h, v = _c128(h_rkb, n_rkb * n_rkb, "h_rkb"), _c128(v_total, nu * n_rkb, "v_total")
        s, x = _c128(s_full, n_rkb * n_rkb, "s_full"), _c128(x_full, n_rkb * nx, "x_full")
        d = _c128(density, n_rkb * n_rkb, "density")
        f, fp = _out(out, (n_rkb, n_rkb), np.complex128)
        self._call("rr_rkb_fock_ri", ctypes.c_int(slot_ll), ctypes.c_int(slot_ss), ctypes.c_int(precision), _sz(nl),
                   _sz(ns), h.ptr, v.ptr, _sz(n_rkb), s.ptr, x.ptr, _sz(x_full.shape[1]), d.ptr, fp)
        return f

    # ---- MO transforms -------------------------------------------------------------------------
    def build_ri_mo_eri(self, slot_ll, slot_ss, nl, ns, c_ukb, precision=FP64, out_precision=FP64, out=None):
        """buildRiMoEri -- src/UKB/RiMoEri.cpp:64-86. Returns (n_aux, n_total, n_total) complex."""
        naux = self.tensor_info(slot_ll)[0]
        nt = c_ukb.shape[1]
        c = _c128(c_ukb, (2 * nl + 2 * ns) * nt, "c_ukb")
        dt = np.complex128 if out_precision == FP64 else np.complex64
        if out is not None:
            _need(out, naux * nt * nt, "out")
            o, op = out, ctypes.c_void_p(out.data_ptr())
        else:
            o = np.empty((naux, nt, nt), dtype=dt)
            op = ctypes.c_void_p(o.ctypes.data)
        self._call("rr_build_ri_mo_eri", ctypes.c_int(slot_ll), ctypes.c_int(slot_ss), ctypes.c_int(precision),
                   ctypes.c_int(out_precision), _sz(nl), _sz(ns), c.ptr, _sz(nt), op)
        return o

    def ri_mo_eri_rotated(self, b, u, precision=FP64, out=None):
        """RiMoEri::rotated -- src/UKB/RiMoEri.cpp:88-110."""
        naux, nt = b.shape[0], u.shape[0]
        bb, uu = _c128(b, naux * nt * nt, "b"), _c128(u, nt * nt, "u")
        o, op = _out(out, (naux, nt, nt), np.complex128)
        self._call("rr_ri_mo_eri_rotated", ctypes.c_int(precision), _sz(nt), _sz(naux), bb.ptr, uu.ptr, op)
        return o

    def ri_mo_eri_positive_block(self, b, off, out=None):
        """RiMoEri::positiveBlock -- src/UKB/RiMoEri.cpp:112-121."""
        naux, nt = b.shape[0], b.shape[1]
        m = nt - off
        bb = _c128(b)
        o, op = _out(out, (naux, m, m), np.complex128)
        self._call("rr_ri_mo_eri_positive_block", _sz(nt), _sz(naux), bb.ptr, _sz(off), op)
        return o

    def build_ri_nonrel_spin_mo_eri(self, slot, c_spatial, precision=FP64, out=None):
        """buildRiNonRelSpinMoEri -- src/NON_REL/NonRelSpinRiMoEri.cpp:13-43."""
        naux, n, _ = self.tensor_info(slot)
        nmo = c_spatial.shape[1]
        c = _f64(c_spatial, n * nmo, "c_spatial")
        o, op = _out(out, (naux, nmo, nmo), np.float64)
        self._call("rr_build_ri_nonrel_spin_mo_eri", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), c.ptr,
                   _sz(nmo), op)
        return o

    def ri_nonrel_spin_mo_eri_rotated(self, b, u, precision=FP64, out=None):
        """RiNonRelSpinMoEri::rotated -- src/NON_REL/NonRelSpinRiMoEri.cpp:45-95."""
        naux, ns = b.shape[0], u.shape[0] // 2
        bb, uu = _f64(b, naux * ns * ns, "b"), _f64(u, 4 * ns * ns, "u")
        o, op = _out(out, (naux, ns, ns), np.float64)
        self._call("rr_ri_nonrel_spin_mo_eri_rotated", ctypes.c_int(precision), _sz(ns), _sz(naux), bb.ptr, uu.ptr, op)
        return o

    def ri_nonrel_spin_embedded_b(self, b, out=None):
        """RiNonRelSpinMoEri::embeddedB -- src/NON_REL/NonRelSpinRiMoEri.h:77-89."""
        naux, ns = b.shape[0], b.shape[1]
        bb = _f64(b)
        o, op = _out(out, (naux, 2 * ns, 2 * ns), np.float64)
        self._call("rr_ri_nonrel_spin_embedded_b", _sz(ns), _sz(naux), bb.ptr, op)
        return o

    # ---- Cholesky ------------------------------------------------------------------------------
    def nonrel_fock_cholesky(self, slot, h_core, p, precision=FP64, out=None):
        """nonRelFockMatrix(h, AoCholesky, P) -- src/Utils/AoCholesky.cpp:108-149."""
        n = h_core.shape[0]
        h, pp = _f64(h_core, n * n, "h_core"), _f64(p, n * n, "p")
        f, fp = _out(out, (n, n), np.float64)
        self._call("rr_nonrel_fock_cholesky", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), h.ptr, pp.ptr, fp)
        return f

    def x2c_fock_cholesky(self, slot, h_x2c, p, precision=FP64, out=None):
        """x2cFockMatrix(h, AoCholesky, P) -- src/Utils/AoCholesky.cpp:151-217."""
        n2 = h_x2c.shape[0]
        h, pp = _c128(h_x2c, n2 * n2, "h_x2c"), _c128(p, n2 * n2, "p")
        f, fp = _out(out, (n2, n2), np.complex128)
        self._call("rr_x2c_fock_cholesky", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n2 // 2), h.ptr, pp.ptr,
                   fp)
        return f

    def cholesky_to_mo_spatial(self, slot, c, precision=FP64, out=None):
        """Spatial block of aoCholeskyToMoSpinOrbital -- src/Utils/AoCholesky.cpp:219-243."""
        nchol, n, _ = self.tensor_info(slot)
        nmo = c.shape[1]
        cc = _f64(c, n * nmo, "c")
        o, op = _out(out, (nchol, nmo, nmo), np.float64)
        self._call("rr_cholesky_to_mo_spatial", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), cc.ptr, _sz(nmo),
                   op)
        return o

    def cholesky_to_mo_spinor(self, slot, c, precision=FP64, out=None):
        """aoCholeskyToMoSpinor -- src/Utils/AoCholesky.cpp:270-295."""
        nchol, n, _ = self.tensor_info(slot)
        nmo = c.shape[1]
        cc = _c128(c, 2 * n * nmo, "c")
        o, op = _out(out, (nchol, nmo, nmo), np.complex128)
        self._call("rr_cholesky_to_mo_spinor", ctypes.c_int(slot), ctypes.c_int(precision), _sz(n), cc.ptr, _sz(nmo),
                   op)
        return o

    # ---- functional stage: MO RI tensors (RiGemmFock / RiSigma*) ---------------------------------
    @staticmethod
    def _elem(b):
        return np.complex128 if _is_complex(b) else np.float64

    def ri_hx_fock_gemm(self, b, h_mo, occupations, two_rdm_h, two_rdm_x, pair_of=None, two_rdm_l1=None,
                        two_rdm_l2=None, out=None):
        """riHxFockGemm -- src/Hessian_opt/RiGemmFock.cpp:53-115 (hartreeExchangeFockMatrix /
        jkOnlyFockMatrix on an RI MO tensor b of shape (n_aux, n, n) or (n_aux, n*n))."""
        cplx = _is_complex(b, h_mo)
        dt = np.complex128 if cplx else np.float64
        n = h_mo.shape[0]
        naux = b.shape[0]
        bb, hh = _Arr(b, dt, naux * n * n, "b"), _Arr(h_mo, dt, n * n, "h_mo")
        oc, h2, x2 = _f64(occupations, n, "occupations"), _f64(two_rdm_h, n * n, "two_rdm_h"), _f64(two_rdm_x, n * n, "two_rdm_x")
        pa = None if pair_of is None else np.ascontiguousarray(np.asarray(pair_of, dtype=np.uint64))
        _need(pa, n, "pair_of")
        l1, l2 = _f64(two_rdm_l1, n * n, "two_rdm_l1"), _f64(two_rdm_l2, n * n, "two_rdm_l2")
        o, op = _out(out, (n, n), dt)
        self._call("rr_ri_hx_fock_gemm", ctypes.c_int(int(cplx)), _sz(n), _sz(naux), bb.ptr, hh.ptr, oc.ptr, h2.ptr,
                   x2.ptr, ctypes.c_void_p(0 if pa is None else pa.ctypes.data), l1.ptr, l2.ptr, op)
        return o

    def ri_sigma_toolkit(self, b, kappa):
        """buildRiSigmaHessianToolkit -- src/Hessian_opt/RiSigmaHessianToolkit.cpp:25-56.
        Returns (kb1, kb2, bdiag, kb1diag, kb2diag) as numpy arrays."""
        dt = self._elem(b)
        naux, n = b.shape[0], kappa.shape[0]
        bb, kk = _Arr(b, dt), _f64(kappa)
        kb1, p1 = _out(None, (naux, n, n), dt)
        kb2, p2 = _out(None, (naux, n, n), dt)
        bd, p3 = _out(None, (naux, n), dt)
        k1d, p4 = _out(None, (naux, n), dt)
        k2d, p5 = _out(None, (naux, n), dt)
        self._call("rr_ri_sigma_toolkit", ctypes.c_int(int(dt == np.complex128)), _sz(n), _sz(naux), bb.ptr, kk.ptr,
                   p1, p2, p3, p4, p5)
        return kb1, kb2, bd, k1d, k2d

    def ri_pair_contract(self, target, sign, n, naux, f, g):
        """pairContract -- src/Hessian_opt/RiSigmaGemm.cpp:76-101. f, g: dicts with optional keys
        s (n_aux, n, n), idx (n, n) uint32, w (n, n), row (n_aux, n), post (n, n). target updated in place."""
        dt = np.complex128 if _is_complex(target) else np.float64
        keep = []

        def mk(d):
            op = PairOp()
            for key in ("s", "idx", "w", "row", "post"):
                v = d.get(key)
                if v is None:
                    continue
                kdt = np.uint32 if key == "idx" else (dt if key in ("s", "row") else np.float64)
                a = np.ascontiguousarray(v, dtype=kdt)
                keep.append(a)
                setattr(op, key, a.ctypes.data)
            return op

        fo, go = mk(f), mk(g)
        t = np.ascontiguousarray(target, dtype=dt)
        self._call("rr_ri_pair_contract", ctypes.c_int(int(dt == np.complex128)), _sz(n), _sz(naux),
                   ctypes.c_double(sign), ctypes.byref(fo), ctypes.byref(go), ctypes.c_void_p(t.ctypes.data))
        target[...] = t
        return target

    def ri_diag_contract(self, target, s, ua=None, ub=None):
        """diagContract -- src/Hessian_opt/RiSigmaGemm.cpp:103-121 (target updated in place)."""
        dt = np.complex128 if _is_complex(target) else np.float64
        n, naux = target.shape[0], s.shape[0]
        ss, a, b_ = _Arr(s, dt), _Arr(ua, dt), _Arr(ub, dt)
        t = np.ascontiguousarray(target, dtype=dt)
        self._call("rr_ri_diag_contract", ctypes.c_int(int(dt == np.complex128)), _sz(n), _sz(naux), ss.ptr, a.ptr,
                   b_.ptr, ctypes.c_void_p(t.ctypes.data))
        target[...] = t
        return target

    def ri_accum_u(self, u, alpha, d, wm):
        """accumU -- src/Hessian_opt/RiSigmaGemm.cpp:123-132 (u updated in place)."""
        dt = np.complex128 if _is_complex(u) else np.float64
        naux, n = u.shape
        dd, ww = _Arr(d, dt), _f64(wm)
        t = np.ascontiguousarray(u, dtype=dt)
        self._call("rr_ri_accum_u", ctypes.c_int(int(dt == np.complex128)), _sz(n), _sz(naux), ctypes.c_double(alpha),
                   dd.ptr, ww.ptr, ctypes.c_void_p(t.ctypes.data))
        u[...] = t
        return u

    def _sigma(self, name, cplx, h_mo, b, occupations, two_rdm_h, two_rdm_x, pairs, v, fock=None, pair_of=None,
               two_rdm_l1=None, two_rdm_l2=None, hx=False):
        dt = np.complex128 if cplx else np.float64
        n, naux = h_mo.shape[0], b.shape[0]
        pp = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1, 2))
        npairs = pp.shape[0]
        vv = _f64(v, npairs * (2 if cplx else 1), "v")
        w = np.zeros(npairs * (2 if cplx else 1))
        hh, bb = _Arr(h_mo, dt, n * n, "h_mo"), _Arr(b, dt, naux * n * n, "b")
        oc, hc, xc = _f64(occupations, n, "occupations"), _f64(two_rdm_h, n * n, "two_rdm_h"), _f64(two_rdm_x, n * n, "two_rdm_x")
        args = [_sz(n), hh.ptr, bb.ptr, _sz(naux), oc.ptr, hc.ptr, xc.ptr]
        if hx:
            ff = _Arr(fock, dt, n * n, "fock")
            pa = None if pair_of is None else np.ascontiguousarray(np.asarray(pair_of, dtype=np.uint64))
            _need(pa, n, "pair_of")
            l1, l2 = _f64(two_rdm_l1, n * n, "two_rdm_l1"), _f64(two_rdm_l2, n * n, "two_rdm_l2")
            args += [ff.ptr, ctypes.c_void_p(0 if pa is None else pa.ctypes.data), l1.ptr, l2.ptr]
        args += [_sz(npairs), ctypes.c_void_p(pp.ctypes.data), vv.ptr, ctypes.c_void_p(w.ctypes.data)]
        self._call(name, *args)
        return w

    def jk_only_sigma_hessian_vector_ri(self, h_mo, b, occupations, two_rdm_h, two_rdm_x, pairs, v):
        """jkOnlySigmaHessianVectorRi -- src/Hessian_opt/JkOnlySigmaHessianVectorRi.cpp:504-523."""
        return self._sigma("rr_jk_only_sigma_hessian_vector_ri", False, h_mo, b, occupations, two_rdm_h, two_rdm_x,
                           pairs, v)

    def jk_only_sigma_joint_hessian_vector_ri(self, h_mo, b, occupations, two_rdm_h, two_rdm_x, pairs, v):
        """jkOnlySigmaJointHessianVectorRi -- JkOnlySigmaHessianVectorRi.cpp:525-573."""
        return self._sigma("rr_jk_only_sigma_joint_hessian_vector_ri", True, h_mo, b, occupations, two_rdm_h, two_rdm_x,
                           pairs, v)

    def hartree_exchange_sigma_hessian_vector_ri(self, h_mo, b, occupations, two_rdm_h, two_rdm_x, fock, pairs, v,
                                                 pair_of=None, two_rdm_l1=None, two_rdm_l2=None):
        """hartreeExchangeSigmaHessianVectorRi -- HartreeExchangeSigmaHessianVectorRi.cpp:442-484."""
        return self._sigma("rr_hartree_exchange_sigma_hessian_vector_ri", False, h_mo, b, occupations, two_rdm_h,
                           two_rdm_x, pairs, v, fock, pair_of, two_rdm_l1, two_rdm_l2, hx=True)

    def hartree_exchange_sigma_joint_hessian_vector_ri(self, h_mo, b, occupations, two_rdm_h, two_rdm_x, fock, pairs,
                                                       v, pair_of=None, two_rdm_l1=None, two_rdm_l2=None):
        """hartreeExchangeSigmaJointHessianVectorRi -- HartreeExchangeSigmaHessianVectorRi.cpp:486-539."""
        return self._sigma("rr_hartree_exchange_sigma_joint_hessian_vector_ri", True, h_mo, b, occupations, two_rdm_h,
                           two_rdm_x, pairs, v, fock, pair_of, two_rdm_l1, two_rdm_l2, hx=True)

    # ---- one-time RI tensor preparation (ThreeCenterIntegrals.cpp) ---------------------------------
    def ri_three_center_tensor(self, raw, metric_x, out=None):
        """riThreeCenterTensor's contraction -- src/AO_ints/ThreeCenterIntegrals.cpp:219-236:
        metric_x^T . raw, raw (n_aux_raw, n*n) -> (n_aux_kept, n*n)."""
        naux_raw = raw.shape[0]
        n2 = int(np.prod(raw.shape[1:]))
        n = int(round(np.sqrt(n2)))
        kept = metric_x.shape[1]
        r, x = _f64(raw), _f64(metric_x, naux_raw * kept, "metric_x")
        o, op = _out(out, (kept, n2), np.float64)
        self._call("rr_ri_three_center_tensor", _sz(n), _sz(naux_raw), r.ptr, x.ptr, _sz(kept), op)
        return o

    def transform_ri_three_center_ao_legs(self, eri3, n_ao, t, out=None):
        """transformRiThreeCenterAoLegs -- src/AO_ints/ThreeCenterIntegrals.cpp:191-217 (t^T B_P t)."""
        naux = eri3.shape[0]
        n_out = t.shape[1]
        e, tt = _f64(eri3, naux * n_ao * n_ao, "eri3"), _f64(t, n_ao * n_out, "t")
        o, op = _out(out, (naux, n_out * n_out), np.float64)
        self._call("rr_transform_ri_three_center_ao_legs", _sz(naux), _sz(n_ao), e.ptr, tt.ptr, _sz(n_out), op)
        return o

    def canonical_orthogonalize(self, s, threshold=1e-10):
        """canonicalOrthogonalize -- src/Linear_Algebra/LinearAlgebra.cpp:267-310. Returns X (n x n_kept)."""
        n = s.shape[0]
        ss = _f64(s, n * n, "s")
        x = np.zeros(n * n)
        kept = _sz()
        self._call("rr_canonical_orthogonalize", _sz(n), ss.ptr, ctypes.c_double(threshold),
                   ctypes.c_void_p(x.ctypes.data), ctypes.byref(kept))
        return x[: n * kept.value].reshape(n, kept.value)

    # ---- fused functional-stage Fock builds (RiJkOnlyFock / RiPnofFock / RiNonRelFullOptFock) -------
    def ri_mo_fock_rkb(self, slot_ll, slot_ss, nl, ns, h_rkb, v_total, c_rkb, occupations, two_rdm_h, two_rdm_x,
                       pair_of=None, two_rdm_l1=None, two_rdm_l2=None, out=None):
        """riJkOnlyFockMatrix / riPnofFockMatrix body -- src/Hessian_opt/RiJkOnlyFock.cpp:14-33."""
        n_rkb, n_mo = c_rkb.shape
        hh, vv = _c128(h_rkb, n_rkb * n_rkb, "h_rkb"), _c128(v_total, (2 * nl + 2 * ns) * n_rkb, "v_total")
        cc = _c128(c_rkb)
        oc, h2 = _f64(occupations, n_mo, "occupations"), _f64(two_rdm_h, n_mo * n_mo, "two_rdm_h")
        x2 = _f64(two_rdm_x, n_mo * n_mo, "two_rdm_x")
        pa = None if pair_of is None else np.ascontiguousarray(np.asarray(pair_of, dtype=np.uint64))
        _need(pa, n_mo, "pair_of")
        l1, l2 = _f64(two_rdm_l1, n_mo * n_mo, "two_rdm_l1"), _f64(two_rdm_l2, n_mo * n_mo, "two_rdm_l2")
        o, op = _out(out, (n_mo, n_mo), np.complex128)
        self._call("rr_ri_mo_fock_rkb", ctypes.c_int(slot_ll), ctypes.c_int(slot_ss), _sz(nl), _sz(ns), hh.ptr, vv.ptr,
                   _sz(n_rkb), cc.ptr, _sz(n_mo), oc.ptr, h2.ptr, x2.ptr,
                   ctypes.c_void_p(0 if pa is None else pa.ctypes.data), l1.ptr, l2.ptr, op)
        return o

    def ri_mo_fock_nonrel(self, slot, h_core_ao, c_current, occupations, two_rdm_h, two_rdm_x, pair_of=None,
                          two_rdm_l1=None, two_rdm_l2=None, out=None):
        """riJkOnlyFockMatrixNonRel / riPnofFockMatrixNonRel body -- RiNonRelFullOptFock.cpp:53-98."""
        ns = h_core_ao.shape[0]
        nt = 2 * ns  # spin orbitals: c_current is n x 2n (RiNonRelFullOptFock.cpp:53-98)
        hh, cc = _f64(h_core_ao, ns * ns, "h_core_ao"), _f64(c_current, ns * nt, "c_current")
        oc, h2 = _f64(occupations, nt, "occupations"), _f64(two_rdm_h, nt * nt, "two_rdm_h")
        x2 = _f64(two_rdm_x, nt * nt, "two_rdm_x")
        pa = None if pair_of is None else np.ascontiguousarray(np.asarray(pair_of, dtype=np.uint64))
        _need(pa, nt, "pair_of")
        l1, l2 = _f64(two_rdm_l1, nt * nt, "two_rdm_l1"), _f64(two_rdm_l2, nt * nt, "two_rdm_l2")
        o, op = _out(out, (2 * ns, 2 * ns), np.float64)
        self._call("rr_ri_mo_fock_nonrel", ctypes.c_int(slot), _sz(ns), hh.ptr, cc.ptr, oc.ptr, h2.ptr, x2.ptr,
                   ctypes.c_void_p(0 if pa is None else pa.ctypes.data), l1.ptr, l2.ptr, op)
        return o

    # ---- 4c RKB Cholesky (RkbCholesky.cpp) -------------------------------------------------------
    def rkb_cholesky_upload(self, set_id, large, small0, small1, cross_fwd, cross_bwd):
        """Upload RkbCholesky vectors (large real, the four small components complex), (n_chol, nl, nl)."""
        nchol = large.shape[0]
        nl = int(round(np.sqrt(np.prod(large.shape[1:]))))
        arrs = [_f64(large), _c128(small0), _c128(small1), _c128(cross_fwd), _c128(cross_bwd)]
        self._call("rr_rkb_cholesky_upload", ctypes.c_int(set_id), _sz(nl), _sz(nchol), *[a.ptr for a in arrs])

    def rkb_cholesky_release(self, set_id):
        self._call("rr_rkb_cholesky_release", ctypes.c_int(set_id))

    def rkb_fock_cholesky(self, set_id, h_rkb, p, out=None):
        """rkbFockMatrix(h_rkb, RkbCholesky, P) -- src/C4_DHF/RkbCholesky.cpp:352-448."""
        n = h_rkb.shape[0]
        hh, pp = _c128(h_rkb), _c128(p)
        o, op = _out(out, (n, n), np.complex128)
        self._call("rr_rkb_fock_cholesky", ctypes.c_int(set_id), _sz(n), hh.ptr, pp.ptr, op)
        return o

    def rkb_fock_cholesky_factor(self, set_id, h_rkb, s_full, x_full, p, out=None):
        """rkbFockMatrix(h_rkb, RkbCholesky, P) with the exchange in factor form (RkbCholesky.cpp:381-432
        through the occupied set of P recovered in the RKB metric, UkbFockMatrixRi.cpp:185-211)."""
        n = h_rkb.shape[0]
        nx = x_full.shape[1]
        hh, pp = _c128(h_rkb, n * n, "h_rkb"), _c128(p, n * n, "density")
        ss, xx = _c128(s_full, n * n, "s_full"), _c128(x_full, n * nx, "x_full")
        o, op = _out(out, (n, n), np.complex128)
        self._call("rr_rkb_fock_cholesky_factor", ctypes.c_int(set_id), _sz(n), hh.ptr, ss.ptr, xx.ptr, _sz(nx), pp.ptr,
                   op)
        return o

    def rkb_cholesky_to_mo(self, set_id, c_mo, n_chol, n_negative=0, threshold=1e-10):
        """rkbCholeskyToMo -- src/C4_DHF/RkbCholesky.cpp:450-523. Returns (n_out, m*m) for n_negative = 0,
        else the recompressed (n_out, n_mo*n_mo)."""
        cc = _c128(c_mo)
        nmo = c_mo.shape[1]
        m = nmo - n_negative
        per = m * m if n_negative == 0 else nmo * nmo
        out = np.zeros((n_chol, per), dtype=np.complex128)
        nout = _sz()
        self._call("rr_rkb_cholesky_to_mo", ctypes.c_int(set_id), cc.ptr, _sz(nmo), _sz(n_negative),
                   ctypes.c_double(threshold), ctypes.c_void_p(out.ctypes.data), _sz(n_chol), ctypes.byref(nout))
        return out[: nout.value]

    # ---- batches of molecules (one HIP-graph launch per batch build) -------------------------------
    def batch_create(self, tensors, copies=COPY_FP64):
        """Keep a list of RI tensors (n_aux, n, n) resident as one batch; returns an opaque handle."""
        arrs = [_f64(t) for t in tensors]
        nm = len(arrs)
        naux = (_sz * nm)(*[t.shape[0] for t in tensors])
        ns = (_sz * nm)(*[int(round(np.sqrt(np.prod(t.shape[1:])))) for t in tensors])
        ptrs = (ctypes.c_void_p * nm)(*[a.ptr.value for a in arrs])
        out = ctypes.c_void_p()
        self._call("rr_batch_create", _sz(nm), naux, ns, ptrs, ctypes.c_uint(copies), ctypes.byref(out))
        return (out, [int(x) for x in ns])

    def batch_destroy(self, batch):
        self._call("rr_batch_destroy", batch[0])

    def batch_jk_factor(self, batch, lfacs, precision=FP64, outs=None):
        """J, K of D_m = L_m L_m^T for every molecule in one graph launch. Returns ([J_m], [K_m], replayed)."""
        handle, ns = batch
        nm = len(ns)
        if len(lfacs) != nm:
            raise ValueError(f"batch_jk_factor: {nm} molecules, {len(lfacs)} factors")
        ls = [_f64(l, n * l.shape[1], f"lfacs[{m}]") for m, (l, n) in enumerate(zip(lfacs, ns))]
        r = (_sz * nm)(*[l.shape[1] for l in lfacs])
        lp = (ctypes.c_void_p * nm)(*[a.ptr.value for a in ls])
        if outs is None:
            js = [np.empty((n, n)) for n in ns]
            ks = [np.empty((n, n)) for n in ns]
        else:
            js, ks = outs
        jp = (ctypes.c_void_p * nm)(*[_Arr(x, np.float64, n * n, "J out").ptr.value for x, n in zip(js, ns)])
        kp = (ctypes.c_void_p * nm)(*[_Arr(x, np.float64, n * n, "K out").ptr.value for x, n in zip(ks, ns)])
        rep = ctypes.c_int(0)
        self._call("rr_batch_jk_factor", handle, ctypes.c_int(precision), r, lp, jp, kp, ctypes.byref(rep))
        return js, ks, bool(rep.value)

    # ---- SCF starting-density network ----------------------------------------------------------------
    def mlp_create(self, n_in, n_hidden, n_layers, n_out, seed=0):
        """FP32 network n_in -> n_layers x (n_hidden, tanh) -> n_out on the device; returns a handle."""
        out = ctypes.c_void_p()
        self._call("rr_mlp_create", ctypes.c_int(n_in), ctypes.c_int(n_hidden), ctypes.c_int(n_layers),
                   ctypes.c_int(n_out), ctypes.c_ulonglong(seed), ctypes.byref(out))
        return out

    def mlp_destroy(self, mlp):
        self._call("rr_mlp_destroy", mlp)

    def mlp_num_params(self, mlp):
        n = _sz()
        self._call("rr_mlp_num_params", mlp, ctypes.byref(n))
        return n.value

    def mlp_get_params(self, mlp):
        p = np.empty(self.mlp_num_params(mlp), dtype=np.float32)
        self._call("rr_mlp_get_params", mlp, ctypes.c_void_p(p.ctypes.data))
        return p

    def mlp_set_params(self, mlp, params):
        a = _Arr(params, np.float32)
        if a.obj.size != self.mlp_num_params(mlp):
            raise ValueError("parameter vector has the wrong length")
        self._call("rr_mlp_set_params", mlp, a.ptr)

    def mlp_forward(self, mlp, x, n_out):
        xa = _f64(x)
        b = x.shape[0]
        y = np.empty((b, n_out))
        self._call("rr_mlp_forward", mlp, _sz(b), xa.ptr, ctypes.c_void_p(y.ctypes.data))
        return y

    def mlp_loss_grad(self, mlp, x, target):
        """Loss 1/(2 B n_out) sum (y - t)^2 and its gradient with respect to the parameter vector."""
        xa, ta = _f64(x), _f64(target)
        g = np.empty(self.mlp_num_params(mlp), dtype=np.float32)
        loss = ctypes.c_double()
        self._call("rr_mlp_loss_grad", mlp, _sz(x.shape[0]), xa.ptr, ta.ptr, ctypes.byref(loss),
                   ctypes.c_void_p(g.ctypes.data))
        return loss.value, g

    def mlp_train_step(self, mlp, x, target, lr):
        """One Adam step on (x, target); returns the loss before the step."""
        xa, ta = _f64(x), _f64(target)
        loss = ctypes.c_double()
        self._call("rr_mlp_train_step", mlp, _sz(x.shape[0]), xa.ptr, ta.ptr, ctypes.c_double(lr),
                   ctypes.byref(loss))
        return loss.value

    # ---- SCF DIIS ------------------------------------------------------------------------------
    def scf_commutator_error(self, fock, density, overlap, out=None):
        """scfCommutatorError -- src/Utils/DIIS.cpp:51-58: F (P S) - (S P) F (real or complex)."""
        cplx = _is_complex(fock, density, overlap)
        conv = _c128 if cplx else _f64
        f, p, s = conv(fock), conv(density), conv(overlap)
        n = fock.shape[0]
        e, ep = _out(out, (n, n), np.complex128 if cplx else np.float64)
        self._call("rr_scf_commutator_error", ctypes.c_int(int(cplx)), _sz(n), f.ptr, p.ptr, s.ptr, ep)
        return e

    def diis_create(self, max_vectors):
        """Diis<T>(max_vectors) -- src/Utils/DIIS.h:50-70, with its history in device memory."""
        out = ctypes.c_void_p()
        self._call("rr_diis_create", _sz(max_vectors), ctypes.byref(out))
        return out

    def diis_destroy(self, diis):
        self._call("rr_diis_destroy", diis)

    def diis_clear(self, diis):
        self._call("rr_diis_clear", diis)

    def diis_extrapolate(self, diis, error, value, out=None):
        """Diis<T>::extrapolate -- src/Utils/DIIS.cpp:80-135. Returns (extrapolated value, weights)."""
        cplx = _is_complex(value, error)
        conv = _c128 if cplx else _f64
        e, v = conv(error), conv(value)
        cw = 2 if cplx else 1
        o, op = _out(out, value.shape, np.complex128 if cplx else np.float64)
        w = np.zeros(16)
        nw = _sz()
        self._call("rr_diis_extrapolate", diis, _sz(cw * int(np.prod(error.shape))), e.ptr,
                   _sz(cw * int(np.prod(value.shape))), v.ptr, op, ctypes.c_void_p(w.ctypes.data), ctypes.byref(nw))
        return o, w[:nw.value].copy()

    # ---- CASSCF RI contractions ----------------------------------------------------------------
    def cas_two_electron_operator(self, b, rho, out=None):
        """casTwoElectronOperator on an RI MO tensor -- src/MCSCF/CasFock.cpp:73-113, :246-259.
        b: (n_aux, n, n) (or (n_aux, n*n)); rho: (n, n) Hermitian. Returns G (n, n)."""
        cplx = _is_complex(b, rho)
        conv = _c128 if cplx else _f64
        n = rho.shape[0]
        naux = b.shape[0]
        bb, rr_ = conv(b), conv(rho)
        g, gp = _out(out, (n, n), np.complex128 if cplx else np.float64)
        self._call("rr_cas_two_electron_operator_ri", ctypes.c_int(int(cplx)), _sz(naux), _sz(n), bb.ptr, rr_.ptr, gp)
        return g

    def cas_active_integrals(self, b, n, active):
        """casActiveIntegrals -- src/MCSCF/CasFock.cpp:136-148: (tu|vw) as an (a^2, a^2) array."""
        cplx = _is_complex(b)
        bb = (_c128 if cplx else _f64)(b)
        act = (ctypes.c_size_t * max(1, len(active)))(*active)
        a2 = len(active) ** 2
        out = np.empty((a2, a2), dtype=np.complex128 if cplx else np.float64)
        self._call("rr_cas_active_integrals_ri", ctypes.c_int(int(cplx)), _sz(b.shape[0]), _sz(n), bb.ptr,
                   _sz(len(active)), act, ctypes.c_void_p(out.ctypes.data))
        return out

    def cas_active_cumulant_fock(self, b, n, active, lam):
        """casActiveCumulantFock -- src/MCSCF/CasFock.cpp:115-134: F (n, a)."""
        cplx = _is_complex(b, lam)
        conv = _c128 if cplx else _f64
        bb, ll = conv(b), conv(lam)
        act = (ctypes.c_size_t * max(1, len(active)))(*active)
        f = np.empty((n, len(active)), dtype=np.complex128 if cplx else np.float64)
        self._call("rr_cas_active_cumulant_fock_ri", ctypes.c_int(int(cplx)), _sz(b.shape[0]), _sz(n), bb.ptr,
                   _sz(len(active)), act, ll.ptr, ctypes.c_void_p(f.ctypes.data))
        return f

    def cc_physics_block(self, b, n, p, q, r, s):
        """FactoredCcIntegrals::physicsBlock over an RI MO tensor -- src/CC/Uccsd.cpp:128-153:
        out[p][q][r][s] = sum_aux B(aux, p n + r) B(aux, q n + s) for the index lists."""
        cplx = _is_complex(b)
        bb = (_c128 if cplx else _f64)(b)
        lists = [(ctypes.c_size_t * max(1, len(x)))(*x) for x in (p, q, r, s)]
        out = np.empty((len(p), len(q), len(r), len(s)), dtype=np.complex128 if cplx else np.float64)
        self._call("rr_cc_physics_block", ctypes.c_int(int(cplx)), _sz(b.shape[0]), _sz(n), bb.ptr,
                   _sz(len(p)), lists[0], _sz(len(q)), lists[1], _sz(len(r)), lists[2], _sz(len(s)), lists[3],
                   ctypes.c_void_p(out.ctypes.data))
        return out
