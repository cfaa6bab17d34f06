// librerdmft_gpu, functional stage: the RI generalized-Fock GEMM of the RDMFT orbital optimiser and
// the RI sigma-vector (orbital-Hessian times vector) engine, on MO-basis RI tensors.
//
// Data: an MO RI tensor B (n_aux, n*n), element B(P,p,q) at [P*n*n + p*n + q], real or complex
// (interleaved re/im), with eri(a,b,c,d) = sum_P B(P,a,c) B(P,b,d) (no conjugation).
//
// Every O(n_aux n^3) contraction is one NT GEMM of rr_gemm.hpp over the flat index k = (P,t):
//   R(x,y) = sum_{P,t} F(P,t,x) G(P,t,y)      with F, G written as F_t[x][k], G_t[y][k]
// by a gather kernel that applies the operand's index map, weights, per-(P,t) row factor and the
// optional right-multiplication ("post", itself a GEMM) on the fly. Rows of every GEMM operand are
// zero-padded to a multiple of 8 elements; the flat k is processed in chunks of aux functions so
// the gathered operands stay within a fixed workspace.
#include <cstdint>
#include <cstdlib>

#include "rr_core.hpp"

namespace rr {
namespace {

enum SigWs {
  WS_SG_F = WS_SIGMA_BASE, WS_SG_G, WS_SG_V, WS_SG_PART, WS_SG_POSTT, WS_SG_WF, WS_SG_WG, WS_SG_IDXF, WS_SG_IDXG,
  WS_SG_B, WS_SG_BT, WS_SG_BP, WS_SG_KB1, WS_SG_KB2, WS_SG_BD, WS_SG_K1D, WS_SG_K2D, WS_SG_KMAT, WS_SG_KMATT,
  WS_SG_H, WS_SG_U1, WS_SG_U2, WS_SG_U3, WS_SG_PHI, WS_SG_PSI, WS_SG_TMP1, WS_SG_TMP2, WS_SG_TMP3, WS_SG_SMALL,
  WS_SG_OUT, WS_SG_TARGET, WS_SG_DPAD, WS_SG_WMP, WS_SG_Y, WS_SG_W, WS_SG_CMAT, WS_SG_PAIR, WS_SG_X2, WS_SG_H2T,
  WS_SG_FOCK, WS_SG_OCC, WS_SG_HC, WS_SG_XC, WS_SG_FOCKOUT, WS_SG_TARGET2, WS_SG_KIN, WS_SG_FOCKIN,
  WS_SG_CUKB, WS_SG_MO, WS_SG_HMO, WS_SG_HTMP, WS_SG_CSP, WS_SG_V2, WS_SG_C2, WS_SG_H2IN,
  WS_SG_OPF = WS_SIGMA_BASE + 500, WS_SG_OPG = WS_SIGMA_BASE + 510
};

// Gathered operand size per chunk of aux functions (logical elements). RR_SIGMA_CHUNK_ELEMS
// overrides it (tests use a small value to exercise the chunk loop; results do not depend on it
// beyond the order of the chunk sums).
size_t chunk_elems() {
  const char* e = std::getenv("RR_SIGMA_CHUNK_ELEMS");
  if (e && *e) {
    const long long v = std::atoll(e);
    if (v > 0) return static_cast<size_t>(v);
  }
  return size_t(1) << 27;
}

// ------------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------------
#define RR_GRID_LOOP(i, total) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (total); i += (size_t)gridDim.x * blockDim.x)

template <bool CPLX>
__device__ inline void load_el(const double* p, size_t e, double& re, double& im) {
  if constexpr (CPLX) {
    re = p[2 * e], im = p[2 * e + 1];
  } else {
    re = p[e], im = 0.0;
  }
}
template <bool CPLX>
__device__ inline void store_el(double* p, size_t e, double re, double im) {
  if constexpr (CPLX) {
    p[2 * e] = re, p[2 * e + 1] = im;
  } else {
    p[e] = re;
  }
}

// Gathered pair-contraction operand for aux functions [p0, p0 + pc):
//   value(P,t,x) = [src ? src(P, idx(t,x)) : 1] * [w ? w(t,x) : 1] * [row ? row(P,t) : 1]
// with idx(t,x) = t*n + x (mode 0), x*n + t (mode 1) or the table idx[t*n + x] (mode 2).
// layout 0: out[x][k], k = (P - p0)*n + t, rows of length ldk (k >= pc*n zero);
// layout 1: out[k][x], ldk rows of length ldx (x >= n or k >= pc*n zero) -- the input of a "post" GEMM.
template <bool CPLX>
__global__ void k_fill_op(const double* src, int mode, const uint32_t* idx, const double* w, const double* row, size_t n,
                          size_t p0, size_t pc, size_t ldk, size_t ldx, int layout, double* out) {
  const size_t n2 = n * n, kk = pc * n;
  const size_t total = layout == 0 ? n * ldk : ldk * ldx;
  RR_GRID_LOOP(i, total) {
    size_t x, k;
    if (layout == 0) {
      x = i / ldk, k = i % ldk;
    } else {
      k = i / ldx, x = i % ldx;
    }
    double re = 0.0, im = 0.0;
    if (k < kk && x < n) {
      const size_t pl = k / n, t = k % n, P = p0 + pl;
      re = 1.0;
      if (src) {
        const size_t j = mode == 0 ? t * n + x : (mode == 1 ? x * n + t : static_cast<size_t>(idx[t * n + x]));
        load_el<CPLX>(src, P * n2 + j, re, im);
      }
      if (w) {
        const double s = w[t * n + x];
        re *= s, im *= s;
      }
      if (row) {
        double rr_, ri;
        load_el<CPLX>(row, P * n + t, rr_, ri);
        const double nr = re * rr_ - im * ri, ni = re * ri + im * rr_;
        re = nr, im = ni;
      }
    }
    store_el<CPLX>(out, i, re, im);
  }
}

// target[x][y] += sign * sum_s part[s][x][y]   (n x n, T-typed, fixed summation order)
template <bool CPLX>
__global__ void k_reduce_into(const double* part, int nsplit, size_t n, double sign, double* target) {
  constexpr int CW = CPLX ? 2 : 1;
  const size_t cnt = n * n * CW;
  RR_GRID_LOOP(i, cnt) {
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += part[z * cnt + i];
    target[i] += sign * s;
  }
}

// Partial sums of target(a,b) += sum_P S(P, b*n+a) (Ua(P,a) + Ub(P,b)), split over P:
//   part[z][b*n + a] over P in [z*pper, (z+1)*pper).
template <bool CPLX>
__global__ void k_diag_part(const double* S, size_t n, size_t naux, const double* Ua, const double* Ub, size_t pper,
                            double* part) {
  const size_t n2 = n * n;
  const size_t z = blockIdx.y;
  const size_t p_beg = z * pper, p_end = min(naux, p_beg + pper);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / n, a = e % n;
    double accr = 0.0, acci = 0.0;
    for (size_t P = p_beg; P < p_end; ++P) {
      double sr, si, ur = 0.0, ui = 0.0, tr, ti;
      load_el<CPLX>(S, P * n2 + e, sr, si);
      if (Ua) {
        load_el<CPLX>(Ua, P * n + a, tr, ti);
        ur += tr, ui += ti;
      }
      if (Ub) {
        load_el<CPLX>(Ub, P * n + b, tr, ti);
        ur += tr, ui += ti;
      }
      accr += sr * ur - si * ui;
      acci += sr * ui + si * ur;
    }
    store_el<CPLX>(part, z * n2 + e, accr, acci);
  }
}
// target[a][b] += sum_z part[z][b][a]
template <bool CPLX>
__global__ void k_diag_reduce(const double* part, int nsplit, size_t n, double* target) {
  const size_t n2 = n * n;
  RR_GRID_LOOP(e, n2) {
    const size_t a = e / n, b = e % n;
    double sr = 0.0, si = 0.0;
    for (int z = 0; z < nsplit; ++z) {
      double r, i;
      load_el<CPLX>(part, z * n2 + b * n + a, r, i);
      sr += r, si += i;
    }
    double tr, ti;
    load_el<CPLX>(target, e, tr, ti);
    store_el<CPLX>(target, e, tr + sr, ti + si);
  }
}

// out[r][c] (row length ldo, zero-padded) = in[r][c] (row length ldi) for r < rows, c < cols;
// trans: out[r][c] = in[c][r]. Real inputs are promoted when the output is complex (in_cw = 1).
template <bool CPLX>
__global__ void k_pad(const double* in, int in_cw, size_t ldi, int trans, size_t rows, size_t cols, size_t ldo,
                      size_t rows_out, double* out) {
  RR_GRID_LOOP(i, rows_out * ldo) {
    const size_t r = i / ldo, c = i % ldo;
    double re = 0.0, im = 0.0;
    if (r < rows && c < cols) {
      const size_t e = trans ? c * ldi + r : r * ldi + c;
      if (in_cw == 2) {
        re = in[2 * e], im = in[2 * e + 1];
      } else {
        re = in[e];
      }
    }
    store_el<CPLX>(out, i, re, im);
  }
}

// Per-aux-row transpose with padding: out[P][b][s] (row length ld) = in[P][s][b]; also the padded
// copy outp[P][a][r] = in[P][a][r] and the diagonal d[P][t] (row length ld) = in[P][t][t].
template <bool CPLX>
__global__ void k_b_layouts(const double* in, size_t n, size_t naux, size_t ld, double* bt, double* bp, double* d) {
  const size_t per = n * ld;
  RR_GRID_LOOP(i, naux * per) {
    const size_t P = i / per, rem = i % per, r = rem / ld, c = rem % ld;
    double tr = 0.0, ti = 0.0, pr = 0.0, pi = 0.0;
    if (c < n) {
      load_el<CPLX>(in, P * n * n + c * n + r, tr, ti);
      load_el<CPLX>(in, P * n * n + r * n + c, pr, pi);
    }
    if (bt) store_el<CPLX>(bt, i, tr, ti);
    if (bp) store_el<CPLX>(bp, i, pr, pi);
    if (d && r == 0) {
      double dr = 0.0, di = 0.0;
      if (c < n) load_el<CPLX>(in, P * n * n + c * n + c, dr, di);
      store_el<CPLX>(d, P * ld + c, dr, di);
    }
  }
}
// d[P][t] (row length ld) = m[P][t][t]
template <bool CPLX>
__global__ void k_diag_of(const double* m, size_t n, size_t naux, size_t ld, double* d) {
  RR_GRID_LOOP(i, naux * ld) {
    const size_t P = i / ld, t = i % ld;
    double re = 0.0, im = 0.0;
    if (t < n) load_el<CPLX>(m, P * n * n + t * n + t, re, im);
    store_el<CPLX>(d, i, re, im);
  }
}

// Y_t[q][k] (k = (P - p0)*n + t, row length ldk) of the Hartree/exchange Fock GEMM:
//   -x2(q,t) B(P,q,t) + [pair] 2 cmat(pair[t], q) B(P, pair[t], pair[q]) + [t == q] W(P,q)
template <bool CPLX>
__global__ void k_hx_y(const double* B, size_t n, size_t p0, size_t pc, size_t ldk, const double* x2, const int64_t* pair,
                       const double* cmat, const double* W, double* Y) {
  const size_t n2 = n * n, kk = pc * n;
  RR_GRID_LOOP(i, n * ldk) {
    const size_t q = i / ldk, k = i % ldk;
    double re = 0.0, im = 0.0;
    if (k < kk) {
      const size_t pl = k / n, t = k % n, P = p0 + pl;
      double br, bi;
      load_el<CPLX>(B, P * n2 + q * n + t, br, bi);
      const double s = -x2[q * n + t];
      re = s * br, im = s * bi;
      if (pair) {
        const size_t st = static_cast<size_t>(pair[t]), qb = static_cast<size_t>(pair[q]);
        load_el<CPLX>(B, P * n2 + st * n + qb, br, bi);
        const double c = 2.0 * cmat[st * n + q];
        re += c * br, im += c * bi;
      }
      if (t == q) {
        load_el<CPLX>(W, P * n + q, br, bi);
        re += br, im += bi;
      }
    }
    store_el<CPLX>(Y, i, re, im);
  }
}

// F(p,q) = occ_q h(q,p) + G(p,q)
template <bool CPLX>
__global__ void k_hx_finish(const double* h, const double* occ, const double* G, size_t n, double* F) {
  RR_GRID_LOOP(e, n * n) {
    const size_t p = e / n, q = e % n;
    double hr, hi, gr, gi;
    load_el<CPLX>(h, q * n + p, hr, hi);
    load_el<CPLX>(G, e, gr, gi);
    store_el<CPLX>(F, e, occ[q] * hr + gr, occ[q] * hi + gi);
  }
}

// Phi(p,q) += (occ_p - occ_q) (KH(q,p) - HK(q,p))
template <bool CPLX>
__global__ void k_phi_comm(const double* KH, const double* HK, const double* occ, size_t n, double* Phi) {
  RR_GRID_LOOP(e, n * n) {
    const size_t p = e / n, q = e % n;
    double ar, ai, br, bi, tr, ti;
    load_el<CPLX>(KH, q * n + p, ar, ai);
    load_el<CPLX>(HK, q * n + p, br, bi);
    load_el<CPLX>(Phi, e, tr, ti);
    const double s = occ[p] - occ[q];
    store_el<CPLX>(Phi, e, tr + s * (ar - br), ti + s * (ai - bi));
  }
}

// Psi(c,d) += sum_a K(a,c)(occ_a - occ_c) h(d,a) - sum_b K(d,b)(occ_d - occ_b) h(b,c)
//          = [h . Ka](d,c) - [Kb . h](d,c) with Ka(a,c) = K(a,c)(occ_a - occ_c), Kb(d,b) = K(d,b)(occ_d - occ_b)
// given HKa = h . Ka and KbH = Kb . h (both n x n, row-major).
template <bool CPLX>
__global__ void k_psi_comm(const double* HKa, const double* KbH, size_t n, double* Psi) {
  RR_GRID_LOOP(e, n * n) {
    const size_t c = e / n, d = e % n;
    double ar, ai, br, bi, tr, ti;
    load_el<CPLX>(HKa, d * n + c, ar, ai);
    load_el<CPLX>(KbH, d * n + c, br, bi);
    load_el<CPLX>(Psi, e, tr, ti);
    store_el<CPLX>(Psi, e, tr + ar - br, ti + ai - bi);
  }
}

// out(r,c) (n x n, T) = in(r,c) * (occ_r - occ_c)   (mode 0) or in(r,c) * (occ_r - occ_c) with the
// occupations of the row minus the column -- both cases share one form; real input promoted.
template <bool CPLX>
__global__ void k_scale_occ_diff(const double* in, const double* occ, size_t n, double* out) {
  RR_GRID_LOOP(e, n * n) {
    const size_t r = e / n, c = e % n;
    store_el<CPLX>(out, e, in[e] * (occ[r] - occ[c]), 0.0);
  }
}

// Hartree/exchange Phi/Psi small block:
//   Phi(a,b) += sum_s K(b,s) F(a,s) + sum_r conj(F(b,r)) K(r,a) - occ_b sum_s K(b,s) h(s,a) - occ_a sum_r h(b,r) K(r,a)
// given FK = F . K^T (FK(a,b) = sum_s F(a,s) K(b,s)), CFK = conj(F) . K  (CFK(b,a) = sum_r conj(F(b,r)) K(r,a)),
// KH = K . h (KH(b,a)), HK = h . K (HK(b,a)).
template <bool CPLX>
__global__ void k_phi_hx_small(const double* FKt, const double* CFK, const double* KH, const double* HK,
                               const double* occ, size_t n, double* Phi) {
  RR_GRID_LOOP(e, n * n) {
    const size_t a = e / n, b = e % n;
    double r0, i0, r1, i1, r2, i2, r3, i3, tr, ti;
    load_el<CPLX>(FKt, a * n + b, r0, i0);
    load_el<CPLX>(CFK, b * n + a, r1, i1);
    load_el<CPLX>(KH, b * n + a, r2, i2);
    load_el<CPLX>(HK, b * n + a, r3, i3);
    load_el<CPLX>(Phi, e, tr, ti);
    store_el<CPLX>(Phi, e, tr + r0 + r1 - occ[b] * r2 - occ[a] * r3, ti + i0 + i1 - occ[b] * i2 - occ[a] * i3);
  }
}
//   Psi(c,d) += sum_a K(a,c) F(a,d) + sum_b K(d,b) conj(F(b,c)) - occ_c sum_a K(a,c) h(d,a) - occ_d sum_b K(d,b) h(b,c)
// given KtF = K^T . F (KtF(c,d)), KCF = K . conj(F) (KCF(d,c)), HK = h . K (HK(d,c)), KH = K . h (KH(d,c)).
template <bool CPLX>
__global__ void k_psi_hx_small(const double* KtF, const double* KCF, const double* HK, const double* KH,
                               const double* occ, size_t n, double* Psi) {
  RR_GRID_LOOP(e, n * n) {
    const size_t c = e / n, d = e % n;
    double r0, i0, r1, i1, r2, i2, r3, i3, tr, ti;
    load_el<CPLX>(KtF, c * n + d, r0, i0);
    load_el<CPLX>(KCF, d * n + c, r1, i1);
    load_el<CPLX>(HK, d * n + c, r2, i2);
    load_el<CPLX>(KH, d * n + c, r3, i3);
    load_el<CPLX>(Psi, e, tr, ti);
    store_el<CPLX>(Psi, e, tr + r0 + r1 - occ[c] * r2 - occ[d] * r3, ti + i0 + i1 - occ[c] * i2 - occ[d] * i3);
  }
}

// h_mo(p,q) = 0 unless p and q are in the same spin block of size ns (buildSpinOneElectron)
__global__ void k_spin_blocks(double* h, size_t ns) {
  const size_t n = 2 * ns;
  RR_GRID_LOOP(e, n * n) {
    if ((e / n) / ns != (e % n) / ns) h[e] = 0.0;
  }
}

template <bool CPLX>
__global__ void k_promote(const double* in, size_t count, double* out) {
  RR_GRID_LOOP(i, count) store_el<CPLX>(out, i, in[i], 0.0);
}
__global__ void k_conj(double* a, size_t count) {
  RR_GRID_LOOP(i, count) a[2 * i + 1] = -a[2 * i + 1];
}
__global__ void k_scale(double* a, size_t count, double s) {
  RR_GRID_LOOP(i, count) a[i] *= s;
}
__global__ void k_fill0(double* a, size_t count) {
  RR_GRID_LOOP(i, count) a[i] = 0.0;
}

// ------------------------------------------------------------------------------------------------
// Host engine
// ------------------------------------------------------------------------------------------------
template <bool CPLX>
constexpr int cw() {
  return CPLX ? 2 : 1;
}

void zero_dev(Context& c, double* p, size_t count) {
  if (count) hipLaunchKernelGGL(k_fill0, dim3(grid_for(count)), dim3(256), 0, c.stream, p, count);
}

// Uploads a host table (or takes a device one) into workspace `id`.
template <typename T>
const T* upload_tab(Context& c, int id, const std::vector<T>& v) {
  T* d = static_cast<T*>(wsbuf(c, id, std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!v.empty()) RR_HIP(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));  // the host vector may go out of scope after return
  return d;
}

// One operand of a pair contraction, all pointers device memory.
struct Op {
  const double* s = nullptr;     // (n_aux, n*n) T, or null (value 1)
  int mode = 0;                  // index map: 0 t*n+x, 1 x*n+t, 2 table
  const uint32_t* idx = nullptr;  // n*n table (mode 2)
  const double* w = nullptr;     // n*n real weights w[t*n+x], or null
  const double* row = nullptr;   // (n_aux, n) T row factors row[P*n+t], or null
  const double* post = nullptr;  // n*n real right factor, or null
};

// Small real-or-complex device product helper (n x n, row-major): C = op(A) op(B).
void small_mm(Context& c, bool cplx, char oa, char ob, size_t n, const double* A, const double* B, double* C) {
  dgemm_small(c, cplx, oa, ob, n, n, n, A, n, B, n, C, n);
}

template <bool CPLX>
struct Engine {
  Context& c;
  size_t n, naux, ldx;
  explicit Engine(Context& cc, size_t nn, size_t na) : c(cc), n(nn), naux(na), ldx(roundup(std::max<size_t>(nn, 1), 8)) {}

  size_t chunk() const { return std::max<size_t>(1, std::min<size_t>(naux, chunk_elems() / std::max<size_t>(1, n * n))); }

  // Gathers op into out[x][k] for aux functions [p0, p0 + pc).
  void fill(const Op& op, size_t p0, size_t pc, size_t ldk, double* out) {
    constexpr int CW = cw<CPLX>();
    if (!op.post) {
      hipLaunchKernelGGL((k_fill_op<CPLX>), dim3(grid_for(n * ldk)), dim3(256), 0, c.stream, op.s, op.mode, op.idx, op.w,
                         op.row, n, p0, pc, ldk, ldx, 0, out);
      RR_HIP(hipGetLastError());
      return;
    }
    // V[k][a] = value(P,t,a); out[x][k] = sum_a post(a,x) V[k][a]  (NT: A = post^T, B = V)
    double* V = static_cast<double*>(wsbuf(c, WS_SG_V, ldk * ldx * CW * sizeof(double)));
    hipLaunchKernelGGL((k_fill_op<CPLX>), dim3(grid_for(ldk * ldx)), dim3(256), 0, c.stream, op.s, op.mode, op.idx, op.w,
                       op.row, n, p0, pc, ldk, ldx, 1, V);
    double* pt = static_cast<double*>(wsbuf(c, WS_SG_POSTT, n * ldx * CW * sizeof(double)));
    hipLaunchKernelGGL((k_pad<CPLX>), dim3(grid_for(n * ldx)), dim3(256), 0, c.stream, op.post, 1, n, 1, n, n, ldx, n, pt);
    RR_HIP(hipGetLastError());
    GemmArgs g;
    g.A = pt, g.lda = (long long)ldx;
    g.B = V, g.ldb = (long long)ldx;
    g.C = out, g.ldcm = (long long)ldk, g.ldcn = 1;
    g.M = (int)n, g.N = (int)ldk, g.K = (int)ldx;
    GemmSpec sp;
    sp.cplx = CPLX, sp.cfg = 1;
    run_gemm(c, sp, g);
  }

  // R = sum over chunks of F_t G_t^T; target += sign * R.
  void pair_contract(double* target, double sign, const Op& f, const Op& g) {
    constexpr int CW = cw<CPLX>();
    if (n == 0 || naux == 0) return;
    const size_t ch = chunk();
    const size_t ldk_max = roundup(ch * n, 8);
    double* Ft = static_cast<double*>(wsbuf(c, WS_SG_F, n * ldk_max * CW * sizeof(double)));
    double* Gt = static_cast<double*>(wsbuf(c, WS_SG_G, n * ldk_max * CW * sizeof(double)));
    for (size_t p0 = 0; p0 < naux; p0 += ch) {
      const size_t pc = std::min(ch, naux - p0);
      const size_t ldk = roundup(pc * n, 8);
      fill(f, p0, pc, ldk, Ft);
      fill(g, p0, pc, ldk, Gt);
      rank_update(target, sign, Ft, Gt, ldk);
    }
  }

  // target[x][y] += sign * sum_k A[x][k] B[y][k]  (n x n, split-K with a fixed-order reduction)
  void rank_update(double* target, double sign, const double* A, const double* B, size_t ldk) {
    constexpr int CW = cw<CPLX>();
    const size_t tiles = ((n + 63) / 64) * ((n + 63) / 64);
    size_t ns = std::max<size_t>(1, std::min<size_t>(std::max<size_t>(1, 1024 / tiles), ldk / 256));
    size_t kper = roundup((ldk + ns - 1) / ns, 16);
    const int nsplit = (int)((ldk + kper - 1) / kper);
    double* part = static_cast<double*>(wsbuf(c, WS_SG_PART, (size_t)nsplit * n * n * CW * sizeof(double)));
    GemmArgs g;
    g.A = A, g.lda = (long long)ldk;
    g.B = B, g.ldb = (long long)ldk;
    g.C = part, g.ldcm = (long long)n, g.ldcn = 1, g.sCz = (long long)(n * n);
    g.M = (int)n, g.N = (int)n, g.K = (int)ldk, g.kper = (int)kper, g.nsplit = nsplit;
    GemmSpec sp;
    sp.cplx = CPLX, sp.cfg = 1;
    gemm_auto(c, sp, g);
    hipLaunchKernelGGL((k_reduce_into<CPLX>), dim3(grid_for(n * n * CW)), dim3(256), 0, c.stream, part, nsplit, n, sign,
                       target);
    RR_HIP(hipGetLastError());
  }

  // target(a,b) += sum_P S(P, b*n+a) (Ua(P,a) + Ub(P,b))
  void diag_contract(double* target, const double* S, const double* Ua, const double* Ub) {
    constexpr int CW = cw<CPLX>();
    if (n == 0 || naux == 0) return;
    const size_t n2 = n * n;
    const size_t bx = std::min<size_t>((n2 + 255) / 256, 4096);
    size_t ns = std::max<size_t>(1, std::min<size_t>(naux, 2048 / bx));
    const size_t pper = (naux + ns - 1) / ns;
    ns = (naux + pper - 1) / pper;
    double* part = static_cast<double*>(wsbuf(c, WS_SG_PART, ns * n2 * CW * sizeof(double)));
    hipLaunchKernelGGL((k_diag_part<CPLX>), dim3((unsigned)bx, (unsigned)ns), dim3(256), 0, c.stream, S, n, naux, Ua, Ub,
                       pper, part);
    hipLaunchKernelGGL((k_diag_reduce<CPLX>), dim3(grid_for(n2)), dim3(256), 0, c.stream, part, (int)ns, n, target);
    RR_HIP(hipGetLastError());
  }

  // U(P,x) += alpha sum_t D[P][t] Wm(x,t); D padded (row length ldx), Wm real n x n, U (n_aux, n).
  void accum_u(double* U, double alpha, const double* Dpad, const double* Wm) {
    constexpr int CW = cw<CPLX>();
    if (n == 0 || naux == 0) return;
    double* wp = static_cast<double*>(wsbuf(c, WS_SG_WMP, n * ldx * CW * sizeof(double)));
    hipLaunchKernelGGL((k_pad<CPLX>), dim3(grid_for(n * ldx)), dim3(256), 0, c.stream, Wm, 1, n, 0, n, n, ldx, n, wp);
    RR_HIP(hipGetLastError());
    GemmArgs g;
    g.A = Dpad, g.lda = (long long)ldx;
    g.B = wp, g.ldb = (long long)ldx;
    g.C = U, g.ldcm = (long long)n, g.ldcn = 1;
    g.M = (int)naux, g.N = (int)n, g.K = (int)ldx, g.alpha = alpha, g.beta = 1.0;
    GemmSpec sp;
    sp.cplx = CPLX, sp.cfg = 1;
    run_gemm(c, sp, g);
  }
};

// Toolkit on device (RiSigmaHessianToolkit.h:44-61): B (caller's, device), KB1[P] = K B_P,
// KB2[P] = B_P K, and the padded diagonals (row length ldx).
struct Toolkit {
  const double* b = nullptr;
  double* kb1 = nullptr;
  double* kb2 = nullptr;
  double* bdiag = nullptr;
  double* kb1diag = nullptr;
  double* kb2diag = nullptr;
};

template <bool CPLX>
Toolkit build_toolkit(Engine<CPLX>& e, const double* dB, const double* dK /* real n x n */) {
  constexpr int CW = cw<CPLX>();
  Context& c = e.c;
  const size_t n = e.n, naux = e.naux, ld = e.ldx, n2 = n * n;
  Toolkit t;
  t.b = dB;
  t.kb1 = static_cast<double*>(wsbuf(c, WS_SG_KB1, std::max<size_t>(1, naux * n2) * CW * sizeof(double)));
  t.kb2 = static_cast<double*>(wsbuf(c, WS_SG_KB2, std::max<size_t>(1, naux * n2) * CW * sizeof(double)));
  t.bdiag = static_cast<double*>(wsbuf(c, WS_SG_BD, std::max<size_t>(1, naux * ld) * CW * sizeof(double)));
  t.kb1diag = static_cast<double*>(wsbuf(c, WS_SG_K1D, std::max<size_t>(1, naux * ld) * CW * sizeof(double)));
  t.kb2diag = static_cast<double*>(wsbuf(c, WS_SG_K2D, std::max<size_t>(1, naux * ld) * CW * sizeof(double)));
  if (n == 0 || naux == 0) return t;
  double* bt = static_cast<double*>(wsbuf(c, WS_SG_BT, naux * n * ld * CW * sizeof(double)));
  double* bp = static_cast<double*>(wsbuf(c, WS_SG_BP, naux * n * ld * CW * sizeof(double)));
  hipLaunchKernelGGL((k_b_layouts<CPLX>), dim3(grid_for(naux * n * ld)), dim3(256), 0, c.stream, dB, n, naux, ld, bt, bp,
                     t.bdiag);
  double* kp = static_cast<double*>(wsbuf(c, WS_SG_KMAT, n * ld * CW * sizeof(double)));
  double* ktp = static_cast<double*>(wsbuf(c, WS_SG_KMATT, n * ld * CW * sizeof(double)));
  hipLaunchKernelGGL((k_pad<CPLX>), dim3(grid_for(n * ld)), dim3(256), 0, c.stream, dK, 1, n, 0, n, n, ld, n, kp);
  hipLaunchKernelGGL((k_pad<CPLX>), dim3(grid_for(n * ld)), dim3(256), 0, c.stream, dK, 1, n, 1, n, n, ld, n, ktp);
  RR_HIP(hipGetLastError());
  {  // KB1[P](a,b) = sum_s K(a,s) B_P(s,b) = sum_s K[a][s] BT_P[b][s]
    GemmArgs g;
    g.A = kp, g.lda = (long long)ld, g.sAb = 0;
    g.B = bt, g.ldb = (long long)ld, g.sBb = (long long)(n * ld);
    g.C = t.kb1, g.ldcm = (long long)n, g.ldcn = 1, g.sCb = (long long)n2;
    g.M = (int)n, g.N = (int)n, g.K = (int)ld, g.nbatch = (int)naux;
    GemmSpec sp;
    sp.cplx = CPLX, sp.cfg = 1;
    gemm_auto(c, sp, g);
  }
  {  // KB2[P](a,b) = sum_r B_P(a,r) K(r,b) = sum_r BP_P[a][r] KT[b][r]
    GemmArgs g;
    g.A = bp, g.lda = (long long)ld, g.sAb = (long long)(n * ld);
    g.B = ktp, g.ldb = (long long)ld, g.sBb = 0;
    g.C = t.kb2, g.ldcm = (long long)n, g.ldcn = 1, g.sCb = (long long)n2;
    g.M = (int)n, g.N = (int)n, g.K = (int)ld, g.nbatch = (int)naux;
    GemmSpec sp;
    sp.cplx = CPLX, sp.cfg = 1;
    gemm_auto(c, sp, g);
  }
  hipLaunchKernelGGL((k_diag_of<CPLX>), dim3(grid_for(naux * ld)), dim3(256), 0, c.stream, t.kb1, n, naux, ld, t.kb1diag);
  hipLaunchKernelGGL((k_diag_of<CPLX>), dim3(grid_for(naux * ld)), dim3(256), 0, c.stream, t.kb2, n, naux, ld, t.kb2diag);
  RR_HIP(hipGetLastError());
  return t;
}

// ------------------------------------------------------------------------------------------------
// Host-side table builders (uploaded once per use; O(n^2))
// ------------------------------------------------------------------------------------------------
struct HostMat {
  size_t n = 0;
  const double* p = nullptr;  // real n x n, row-major, host
  double operator()(size_t r, size_t c) const { return p[r * n + c]; }
};

template <typename F>
std::vector<double> wtab(size_t n, F f) {
  std::vector<double> v(n * n);
  for (size_t t = 0; t < n; ++t)
    for (size_t x = 0; x < n; ++x) v[t * n + x] = f(t, x);
  return v;
}
template <typename F>
std::vector<uint32_t> itab(size_t n, F f) {
  std::vector<uint32_t> v(n * n);
  for (size_t t = 0; t < n; ++t)
    for (size_t x = 0; x < n; ++x) v[t * n + x] = static_cast<uint32_t>(f(t, x));
  return v;
}

// A pair-contraction operand built from host descriptions; owns its device tables.
struct OpBuilder {
  Context& c;
  std::vector<std::pair<void*, size_t>> owned;
  explicit OpBuilder(Context& cc) : c(cc) {}
  ~OpBuilder() {
    if (!owned.empty()) (void)hipStreamSynchronize(c.stream);
    for (auto& p : owned) (void)hipFree(p.first);
  }
  template <typename T>
  const T* dev(const std::vector<T>& v) {
    void* p = nullptr;
    RR_HIP(hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)));
    owned.push_back({p, v.size()});
    if (!v.empty()) RR_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return static_cast<const T*>(p);
  }
};

// ------------------------------------------------------------------------------------------------
// riHxFockGemm (RiGemmFock.cpp:53-115)
// ------------------------------------------------------------------------------------------------
template <bool CPLX>
void hx_fock(Context& c, size_t n, size_t naux, const double* dB, const double* dh, const double* docc, const double* dh2,
             const double* dx2, const int64_t* dpair, const double* dcmat, double* dF) {
  constexpr int CW = cw<CPLX>();
  Engine<CPLX> e(c, n, naux);
  const size_t ld = e.ldx;
  double* G = static_cast<double*>(wsbuf(c, WS_SG_TARGET, std::max<size_t>(1, n * n) * CW * sizeof(double)));
  zero_dev(c, G, n * n * CW);
  if (n > 0 && naux > 0) {
    // W(P,q) = sum_s Bdiag(P,s) h2(s,q)
    double* bd = static_cast<double*>(wsbuf(c, WS_SG_BD, naux * ld * CW * sizeof(double)));
    hipLaunchKernelGGL((k_diag_of<CPLX>), dim3(grid_for(naux * ld)), dim3(256), 0, c.stream, dB, n, naux, ld, bd);
    double* W = static_cast<double*>(wsbuf(c, WS_SG_W, naux * n * CW * sizeof(double)));
    zero_dev(c, W, naux * n * CW);
    // accum_u computes U(P,x) += sum_t D(P,t) Wm(x,t); with Wm = h2^T: sum_s Bdiag(P,s) h2(s,x).
    double* h2t = static_cast<double*>(wsbuf(c, WS_SG_H2T, n * n * sizeof(double)));
    hipLaunchKernelGGL((k_pad<false>), dim3(grid_for(n * n)), dim3(256), 0, c.stream, dh2, 1, n, 1, n, n, n, n, h2t);
    e.accum_u(W, 1.0, bd, h2t);
    const size_t ch = e.chunk();
    const size_t ldk_max = roundup(ch * n, 8);
    double* At = static_cast<double*>(wsbuf(c, WS_SG_F, n * ldk_max * CW * sizeof(double)));
    double* Yt = static_cast<double*>(wsbuf(c, WS_SG_Y, n * ldk_max * CW * sizeof(double)));
    for (size_t p0 = 0; p0 < naux; p0 += ch) {
      const size_t pc = std::min(ch, naux - p0);
      const size_t ldk = roundup(pc * n, 8);
      // At[p][k] = B(P,t,p): operand with idx(t,p) = t*n + p
      Op a;
      a.s = dB, a.mode = 0;
      e.fill(a, p0, pc, ldk, At);
      hipLaunchKernelGGL((k_hx_y<CPLX>), dim3(grid_for(n * ldk)), dim3(256), 0, c.stream, dB, n, p0, pc, ldk, dx2, dpair,
                         dcmat, W, Yt);
      RR_HIP(hipGetLastError());
      e.rank_update(G, 1.0, At, Yt, ldk);
    }
  }
  hipLaunchKernelGGL((k_hx_finish<CPLX>), dim3(grid_for(n * n)), dim3(256), 0, c.stream, dh, docc, G, n, dF);
  RR_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------------
// Sigma drivers: Phi/Psi of the JK-only functional (JkOnlySigmaHessianVectorRi.cpp:337-500) and of
// the Hartree/exchange(+L1/L2) functional (HartreeExchangeSigmaHessianVectorRi.cpp:220-439).
// ------------------------------------------------------------------------------------------------
struct SigmaInputs {
  size_t n = 0, naux = 0;
  const double* dB = nullptr;    // device (n_aux, n^2) T
  const double* dh = nullptr;    // device n x n T
  const double* docc = nullptr;  // device n
  HostMat hc, xc;                // host real couplings (two_rdm_h / two_rdm_x)
  const double* dhc = nullptr;   // device copies
  const double* dxc = nullptr;
  const double* dfock = nullptr;  // device n x n T (Hartree/exchange only)
  const std::vector<size_t>* pair_of = nullptr;  // host, empty/null: no L1/L2
  HostMat l1, l2;
};

// Device n x n real matrix (transposed copy helper)
const double* dev_transpose(Context& c, int id, const double* d, size_t n) {
  double* t = static_cast<double*>(wsbuf(c, id, std::max<size_t>(1, n * n) * sizeof(double)));
  hipLaunchKernelGGL((k_pad<false>), dim3(grid_for(n * n)), dim3(256), 0, c.stream, d, 1, n, 1, n, n, n, n, t);
  RR_HIP(hipGetLastError());
  return t;
}

template <bool CPLX>
struct SigmaRun {
  Engine<CPLX>& e;
  const SigmaInputs& in;
  const Toolkit& t;
  const double* dK;    // device real K
  const HostMat& K;    // host real K
  OpBuilder ob;
  SigmaRun(Engine<CPLX>& ee, const SigmaInputs& ii, const Toolkit& tt, const double* dk, const HostMat& hk)
      : e(ee), in(ii), t(tt), dK(dk), K(hk), ob(ee.c) {}

  size_t n() const { return in.n; }
  Op op(const double* s, int mode) {
    Op o;
    o.s = s, o.mode = mode;
    return o;
  }
  template <typename F>
  Op opw(const double* s, int mode, F wf) {
    Op o = op(s, mode);
    o.w = ob.dev(wtab(n(), wf));
    return o;
  }
  template <typename FI>
  Op opi(const double* s, FI fi) {
    Op o;
    o.s = s, o.mode = 2;
    o.idx = ob.dev(itab(n(), fi));
    return o;
  }
  template <typename FI, typename F>
  Op opiw(const double* s, FI fi, F wf) {
    Op o = opi(s, fi);
    o.w = ob.dev(wtab(n(), wf));
    return o;
  }
  double* umat(int id) {
    constexpr int CW = cw<CPLX>();
    double* u = static_cast<double*>(wsbuf(e.c, id, std::max<size_t>(1, in.naux * n()) * CW * sizeof(double)));
    zero_dev(e.c, u, in.naux * n() * CW);
    return u;
  }
  void negate_copy(const double* src, double* dst) {
    constexpr int CW = cw<CPLX>();
    const size_t cnt = in.naux * n() * CW;
    RR_HIP(hipMemcpyAsync(dst, src, cnt * sizeof(double), hipMemcpyDeviceToDevice, e.c.stream));
    if (cnt) hipLaunchKernelGGL(k_scale, dim3(grid_for(cnt)), dim3(256), 0, e.c.stream, dst, cnt, -1.0);
  }
  // K promoted to T (device)
  const double* kT(int id) {
    constexpr int CW = cw<CPLX>();
    double* k = static_cast<double*>(wsbuf(e.c, id, std::max<size_t>(1, n() * n()) * CW * sizeof(double)));
    hipLaunchKernelGGL((k_promote<CPLX>), dim3(grid_for(n() * n())), dim3(256), 0, e.c.stream, dK, n() * n(), k);
    return k;
  }

  // ---- JK-only -------------------------------------------------------------------------------
  void jk_phi(double* Phi) {
    using Z = size_t;
    const size_t nn = n();
    const HostMat& xc = in.xc;
    const double* Bm = t.b;
    const double* K1 = t.kb1;
    const double* K2 = t.kb2;
    {  // T12: commutator
      const double* kt = kT(WS_SG_TMP3);
      double* KH = static_cast<double*>(wsbuf(e.c, WS_SG_TMP1, nn * nn * cw<CPLX>() * sizeof(double)));
      double* HK = static_cast<double*>(wsbuf(e.c, WS_SG_TMP2, nn * nn * cw<CPLX>() * sizeof(double)));
      small_mm(e.c, CPLX, 'N', 'N', nn, kt, in.dh, KH);
      small_mm(e.c, CPLX, 'N', 'N', nn, in.dh, kt, HK);
      hipLaunchKernelGGL((k_phi_comm<CPLX>), dim3(grid_for(nn * nn)), dim3(256), 0, e.c.stream, KH, HK, in.docc, nn, Phi);
    }
    double* Uh = umat(WS_SG_U1);
    e.accum_u(Uh, 1.0, t.bdiag, in.dhc);
    double* nUh = umat(WS_SG_U2);
    negate_copy(Uh, nUh);
    e.diag_contract(Phi, K1, Uh, nUh);
    e.pair_contract(Phi, -1.0, opw(Bm, 0, [&](Z i, Z p) { return xc(p, i); }), op(K1, 1));
    e.pair_contract(Phi, 1.0, op(Bm, 0), opw(K1, 1, [&](Z i, Z q) { return xc(q, i); }));
    e.diag_contract(Phi, K2, nUh, Uh);
    e.pair_contract(Phi, 1.0, opw(K2, 0, [&](Z i, Z p) { return xc(p, i); }), op(Bm, 1));
    e.pair_contract(Phi, -1.0, op(K2, 0), opw(Bm, 1, [&](Z i, Z q) { return xc(q, i); }));
    {  // TA
      double* V = umat(WS_SG_U3);
      e.accum_u(V, 1.0, t.kb1diag, in.dhc);
      e.accum_u(V, -1.0, t.kb2diag, in.dhc);
      double* nV = umat(WS_SG_U2);
      negate_copy(V, nV);
      e.diag_contract(Phi, Bm, V, nV);
    }
    e.pair_contract(Phi, 1.0, opw(Bm, 0, [&](Z i, Z p) { return xc(p, i); }), op(K2, 1));
    e.pair_contract(Phi, -1.0, op(Bm, 0), opw(K2, 1, [&](Z i, Z q) { return xc(q, i); }));
    e.pair_contract(Phi, -1.0, opw(K1, 0, [&](Z i, Z p) { return xc(p, i); }), op(Bm, 1));
    e.pair_contract(Phi, 1.0, op(K1, 0), opw(Bm, 1, [&](Z i, Z q) { return xc(q, i); }));
  }

  void jk_psi(double* Psi) {
    using Z = size_t;
    const size_t nn = n();
    const HostMat& xc = in.xc;
    const double* Bm = t.b;
    const double* K1 = t.kb1;
    const double* K2 = t.kb2;
    {  // Psi(c,d) += sum_a K(a,c)(occ_a - occ_c) h(d,a) - sum_b K(d,b)(occ_d - occ_b) h(b,c)
      double* ko = static_cast<double*>(wsbuf(e.c, WS_SG_TMP3, nn * nn * cw<CPLX>() * sizeof(double)));
      hipLaunchKernelGGL((k_scale_occ_diff<CPLX>), dim3(grid_for(nn * nn)), dim3(256), 0, e.c.stream, dK, in.docc, nn, ko);
      double* HKa = static_cast<double*>(wsbuf(e.c, WS_SG_TMP1, nn * nn * cw<CPLX>() * sizeof(double)));
      double* KbH = static_cast<double*>(wsbuf(e.c, WS_SG_TMP2, nn * nn * cw<CPLX>() * sizeof(double)));
      small_mm(e.c, CPLX, 'N', 'N', nn, in.dh, ko, HKa);   // (h . Ka)(d,c) = sum_a h(d,a) K(a,c)(occ_a - occ_c)
      small_mm(e.c, CPLX, 'N', 'N', nn, ko, in.dh, KbH);   // (Kb . h)(d,c) = sum_b K(d,b)(occ_d - occ_b) h(b,c)
      hipLaunchKernelGGL((k_psi_comm<CPLX>), dim3(grid_for(nn * nn)), dim3(256), 0, e.c.stream, HKa, KbH, nn, Psi);
    }
    double* Uh = umat(WS_SG_U1);
    e.accum_u(Uh, 1.0, t.bdiag, in.dhc);
    double* nUh = umat(WS_SG_U2);
    negate_copy(Uh, nUh);
    {  // P3 (i)
      Op f;
      f.w = ob.dev(wtab(nn, [&](Z a, Z c) { return K(a, c); }));
      f.row = Uh;
      e.pair_contract(Psi, 1.0, f, op(Bm, 1));
    }
    e.diag_contract(Psi, K2, nUh, nullptr);  // P3 (ii)
    {  // P3 (iii)
      Op f = opw(Bm, 0, [&](Z i, Z a) { return xc(a, i); });
      f.post = dK;
      e.pair_contract(Psi, -1.0, f, op(Bm, 1));
    }
    e.pair_contract(Psi, 1.0, opw(K2, 0, [&](Z i, Z c) { return xc(c, i); }), op(Bm, 1));  // P3 (iv)
    e.diag_contract(Psi, K1, nullptr, nUh);                                                  // P4 (i)
    {  // P4 (ii)
      Op f = op(Bm, 0);
      f.row = Uh;
      Op g;
      g.w = ob.dev(wtab(nn, [&](Z b, Z d) { return K(d, b); }));
      e.pair_contract(Psi, 1.0, f, g);
    }
    e.pair_contract(Psi, 1.0, op(Bm, 0), opw(K1, 1, [&](Z i, Z d) { return xc(d, i); }));  // P4 (iii)
    {  // P4 (iv)
      Op g = opw(Bm, 1, [&](Z i, Z b) { return xc(b, i); });
      g.post = dev_transpose(e.c, WS_SG_SMALL, dK, nn);
      e.pair_contract(Psi, -1.0, op(Bm, 0), g);
    }
    {  // P5
      const double* hcT = dev_transpose(e.c, WS_SG_TMP1, in.dhc, nn);
      double* V = umat(WS_SG_U3);
      e.accum_u(V, 1.0, t.kb1diag, hcT);
      e.accum_u(V, -1.0, t.kb2diag, hcT);
      double* nV = umat(WS_SG_U2);
      negate_copy(V, nV);
      e.diag_contract(Psi, Bm, V, nV);
    }
    // P6
    e.pair_contract(Psi, 1.0, op(K1, 0), opw(Bm, 1, [&](Z i, Z d) { return xc(i, d); }));
    e.pair_contract(Psi, -1.0, opw(K1, 0, [&](Z i, Z c) { return xc(i, c); }), op(Bm, 1));
    e.pair_contract(Psi, 1.0, opw(Bm, 0, [&](Z i, Z c) { return xc(i, c); }), op(K2, 1));
    e.pair_contract(Psi, -1.0, op(Bm, 0), opw(K2, 1, [&](Z i, Z d) { return xc(i, d); }));
  }

  // ---- Hartree/exchange (+ L1/L2) ----------------------------------------------------------------
  void hx_phi(double* Phi) {
    using Z = size_t;
    const size_t nn = n();
    const HostMat& hcpl = in.hc;
    const HostMat& xcpl = in.xc;
    const double* Bm = t.b;
    const double* K1 = t.kb1;
    const double* K2 = t.kb2;
    constexpr int CW = cw<CPLX>();
    {
      const double* kt = kT(WS_SG_TMP3);
      double* m0 = static_cast<double*>(wsbuf(e.c, WS_SG_TMP1, nn * nn * CW * sizeof(double)));
      double* m1 = static_cast<double*>(wsbuf(e.c, WS_SG_TMP2, nn * nn * CW * sizeof(double)));
      double* m2 = static_cast<double*>(wsbuf(e.c, WS_SG_SMALL, nn * nn * CW * sizeof(double)));
      double* m3 = static_cast<double*>(wsbuf(e.c, WS_SG_OUT, nn * nn * CW * sizeof(double)));
      small_mm(e.c, CPLX, 'N', 'T', nn, in.dfock, kt, m0);                       // FKt(a,b) = sum_s F(a,s) K(b,s)
      if (CPLX) {
        double* fc = static_cast<double*>(wsbuf(e.c, WS_SG_FOCK, nn * nn * CW * sizeof(double)));
        RR_HIP(hipMemcpyAsync(fc, in.dfock, nn * nn * CW * sizeof(double), hipMemcpyDeviceToDevice, e.c.stream));
        hipLaunchKernelGGL(k_conj, dim3(grid_for(nn * nn)), dim3(256), 0, e.c.stream, fc, nn * nn);
        small_mm(e.c, CPLX, 'N', 'N', nn, fc, kt, m1);                          // CFK(b,a) = sum_r conj(F(b,r)) K(r,a)
      } else {
        small_mm(e.c, CPLX, 'N', 'N', nn, in.dfock, kt, m1);
      }
      small_mm(e.c, CPLX, 'N', 'N', nn, kt, in.dh, m2);                          // KH(b,a) = sum_s K(b,s) h(s,a)
      small_mm(e.c, CPLX, 'N', 'N', nn, in.dh, kt, m3);                          // HK(b,a) = sum_r h(b,r) K(r,a)
      hipLaunchKernelGGL((k_phi_hx_small<CPLX>), dim3(grid_for(nn * nn)), dim3(256), 0, e.c.stream, m0, m1, m2, m3,
                         in.docc, nn, Phi);
      RR_HIP(hipGetLastError());
    }
    double* Uh = umat(WS_SG_U1);
    e.accum_u(Uh, 1.0, t.bdiag, in.dhc);
    double* nUh = umat(WS_SG_U2);
    negate_copy(Uh, nUh);
    e.pair_contract(Phi, 1.0, op(Bm, 0), opw(K1, 1, [&](Z u, Z b) { return hcpl(b, u); }));  // Ea
    e.diag_contract(Phi, K1, nullptr, nUh);                                                   // -Eb
    e.pair_contract(Phi, 1.0, opw(K2, 0, [&](Z u, Z a) { return hcpl(a, u); }), op(Bm, 1));  // Fa
    e.diag_contract(Phi, K2, nUh, nullptr);                                                   // -Fb
    e.pair_contract(Phi, 1.0, op(K1, 0), opw(Bm, 1, [&](Z r, Z b) { return xcpl(r, b); }));  // G1
    e.pair_contract(Phi, -1.0, opw(K1, 0, [&](Z r, Z a) { return xcpl(r, a); }), op(Bm, 1));
    e.pair_contract(Phi, 1.0, opw(Bm, 0, [&](Z s, Z a) { return xcpl(a, s); }), op(K2, 1));  // G2
    e.pair_contract(Phi, -1.0, op(Bm, 0), opw(K2, 1, [&](Z s, Z b) { return hcpl(s, b); }));
    {  // H1/H2
      const double* hcT = dev_transpose(e.c, WS_SG_TMP1, in.dhc, nn);
      const double* xcT = dev_transpose(e.c, WS_SG_TMP2, in.dxc, nn);
      double* Ua = umat(WS_SG_U3);
      double* Ub = umat(WS_SG_U2);
      e.accum_u(Ua, 1.0, t.kb1diag, hcT);
      e.accum_u(Ua, -1.0, t.kb2diag, in.dxc);
      e.accum_u(Ub, -1.0, t.kb1diag, xcT);
      e.accum_u(Ub, 1.0, t.kb2diag, in.dxc);
      e.diag_contract(Phi, Bm, Ua, Ub);
    }
  }

  void hx_psi(double* Psi) {
    using Z = size_t;
    const size_t nn = n();
    const HostMat& hcpl = in.hc;
    const HostMat& xcpl = in.xc;
    const double* Bm = t.b;
    const double* K1 = t.kb1;
    const double* K2 = t.kb2;
    constexpr int CW = cw<CPLX>();
    {
      const double* kt = kT(WS_SG_TMP3);
      double* m0 = static_cast<double*>(wsbuf(e.c, WS_SG_TMP1, nn * nn * CW * sizeof(double)));
      double* m1 = static_cast<double*>(wsbuf(e.c, WS_SG_TMP2, nn * nn * CW * sizeof(double)));
      double* m2 = static_cast<double*>(wsbuf(e.c, WS_SG_SMALL, nn * nn * CW * sizeof(double)));
      double* m3 = static_cast<double*>(wsbuf(e.c, WS_SG_OUT, nn * nn * CW * sizeof(double)));
      small_mm(e.c, CPLX, 'T', 'N', nn, kt, in.dfock, m0);  // KtF(c,d) = sum_a K(a,c) F(a,d)
      if (CPLX) {
        double* fc = static_cast<double*>(wsbuf(e.c, WS_SG_FOCK, nn * nn * CW * sizeof(double)));
        RR_HIP(hipMemcpyAsync(fc, in.dfock, nn * nn * CW * sizeof(double), hipMemcpyDeviceToDevice, e.c.stream));
        hipLaunchKernelGGL(k_conj, dim3(grid_for(nn * nn)), dim3(256), 0, e.c.stream, fc, nn * nn);
        small_mm(e.c, CPLX, 'N', 'N', nn, kt, fc, m1);  // KCF(d,c) = sum_b K(d,b) conj(F(b,c))
      } else {
        small_mm(e.c, CPLX, 'N', 'N', nn, kt, in.dfock, m1);
      }
      small_mm(e.c, CPLX, 'N', 'N', nn, in.dh, kt, m2);  // HK(d,c) = sum_a h(d,a) K(a,c)
      small_mm(e.c, CPLX, 'N', 'N', nn, kt, in.dh, m3);  // KH(d,c) = sum_b K(d,b) h(b,c)
      hipLaunchKernelGGL((k_psi_hx_small<CPLX>), dim3(grid_for(nn * nn)), dim3(256), 0, e.c.stream, m0, m1, m2, m3,
                         in.docc, nn, Psi);
      RR_HIP(hipGetLastError());
    }
    double* Uh = umat(WS_SG_U1);
    e.accum_u(Uh, 1.0, t.bdiag, in.dhc);
    double* nUh = umat(WS_SG_U2);
    negate_copy(Uh, nUh);
    e.pair_contract(Psi, 1.0, opw(K2, 0, [&](Z u, Z c) { return hcpl(c, u); }), op(Bm, 1));  // Ra
    e.diag_contract(Psi, K2, nUh, nullptr);                                                   // -Rb
    e.pair_contract(Psi, 1.0, op(Bm, 0), opw(K1, 1, [&](Z u, Z d) { return hcpl(d, u); }));  // Sa
    e.diag_contract(Psi, K1, nullptr, nUh);                                                   // -Sb
    e.pair_contract(Psi, 1.0, opw(Bm, 0, [&](Z b, Z c) { return xcpl(c, b); }), op(K2, 1));  // U1
    e.pair_contract(Psi, -1.0, op(Bm, 0), opw(K2, 1, [&](Z b, Z d) { return hcpl(d, b); }));
    e.pair_contract(Psi, 1.0, op(K1, 0), opw(Bm, 1, [&](Z a, Z d) { return xcpl(a, d); }));  // U2
    e.pair_contract(Psi, -1.0, opw(K1, 0, [&](Z a, Z c) { return xcpl(c, a); }), op(Bm, 1));
    {  // V1/V2
      const double* xcT = dev_transpose(e.c, WS_SG_TMP2, in.dxc, nn);
      double* Ua = umat(WS_SG_U3);
      double* Ub = umat(WS_SG_U2);
      e.accum_u(Ua, 1.0, t.kb1diag, in.dhc);
      e.accum_u(Ua, -1.0, t.kb2diag, in.dxc);
      e.accum_u(Ub, -1.0, t.kb1diag, xcT);
      e.accum_u(Ub, 1.0, t.kb2diag, xcT);
      e.diag_contract(Psi, Bm, Ua, Ub);
    }
  }

  void l12_phi(double* Phi) {
    using Z = size_t;
    const size_t nn = n();
    const std::vector<size_t>& pr = *in.pair_of;
    const HostMat& l1 = in.l1;
    const HostMat& l2 = in.l2;
    const double* Bm = t.b;
    const double* K1 = t.kb1;
    const double* K2 = t.kb2;
    // I
    e.pair_contract(Phi, -1.0, op(Bm, 0),
                    opiw(K2, [&](Z u, Z b) { return pr[u] * nn + pr[b]; },
                         [&](Z u, Z b) { return 2.0 * (l1(u, pr[b]) + l2(u, b)); }));
    // Jr, Jr2
    e.pair_contract(Phi, 1.0, opi(Bm, [&](Z r, Z a) { return pr[r] * nn + a; }),
                    opiw(K1, [&](Z r, Z b) { return r * nn + pr[b]; },
                         [&](Z r, Z b) { return 2.0 * (l1(r, b) + l2(r, pr[b])); }));
    e.pair_contract(Phi, -1.0, op(K1, 0),
                    opiw(Bm, [&](Z r, Z b) { return pr[r] * nn + pr[b]; },
                         [&](Z r, Z b) { return 2.0 * (l1(r, b) + l2(r, pr[b])); }));
    // Ls, Ls2
    e.pair_contract(Phi, 1.0,
                    opiw(K2, [&](Z s, Z a) { return pr[a] * nn + s; },
                         [&](Z s, Z a) { return 2.0 * (l1(a, s) + l2(a, pr[s])); }),
                    opi(Bm, [&](Z s, Z b) { return b * nn + pr[s]; }));
    e.pair_contract(Phi, -1.0,
                    opiw(Bm, [&](Z s, Z a) { return pr[a] * nn + pr[s]; },
                         [&](Z s, Z a) { return 2.0 * (l1(a, s) + l2(a, pr[s])); }),
                    op(K2, 1));
    // Mw
    e.pair_contract(Phi, 1.0,
                    opiw(K1, [&](Z w, Z a) { return pr[a] * nn + w; },
                         [&](Z w, Z a) { return 2.0 * (l1(pr[a], w) + l2(pr[a], pr[w])); }),
                    opi(Bm, [&](Z w, Z b) { return b * nn + pr[w]; }));
  }

  void l12_psi(double* Psi) {
    using Z = size_t;
    const size_t nn = n();
    const std::vector<size_t>& pr = *in.pair_of;
    const HostMat& l1 = in.l1;
    const HostMat& l2 = in.l2;
    const double* Bm = t.b;
    const double* K1 = t.kb1;
    const double* K2 = t.kb2;
    // Wu
    e.pair_contract(Psi, -1.0, opi(Bm, [&](Z u, Z c) { return pr[u] * nn + c; }),
                    opiw(K2, [&](Z u, Z d) { return u * nn + pr[d]; },
                         [&](Z u, Z d) { return 2.0 * (l1(u, d) + l2(u, pr[d])); }));
    // Xb, Xb2
    e.pair_contract(Psi, 1.0,
                    opiw(K2, [&](Z b, Z c) { return pr[c] * nn + b; },
                         [&](Z b, Z c) { return 2.0 * (l1(c, b) + l2(c, pr[b])); }),
                    opi(Bm, [&](Z b, Z d) { return d * nn + pr[b]; }));
    e.pair_contract(Psi, -1.0,
                    opiw(Bm, [&](Z b, Z c) { return pr[c] * nn + pr[b]; },
                         [&](Z b, Z c) { return 2.0 * (l1(c, b) + l2(c, pr[b])); }),
                    op(K2, 1));
    // Ya, Ya2
    e.pair_contract(Psi, 1.0, opi(Bm, [&](Z a, Z c) { return pr[a] * nn + c; }),
                    opiw(K1, [&](Z a, Z d) { return a * nn + pr[d]; },
                         [&](Z a, Z d) { return 2.0 * (l1(a, d) + l2(a, pr[d])); }));
    e.pair_contract(Psi, -1.0, op(K1, 0),
                    opiw(Bm, [&](Z a, Z d) { return pr[a] * nn + pr[d]; },
                         [&](Z a, Z d) { return 2.0 * (l1(a, d) + l2(a, pr[d])); }));
    // Zw
    e.pair_contract(Psi, 1.0,
                    opiw(K1, [&](Z w, Z c) { return pr[c] * nn + pr[w]; },
                         [&](Z w, Z c) { return 2.0 * (l1(c, w) + l2(c, pr[w])); }),
                    op(Bm, 1));
  }
};

// Phi and Psi (device, n x n T) for one kappa matrix.
template <bool CPLX>
void phi_psi(Context& c, const SigmaInputs& in, bool hx, const std::vector<double>& Kh, double* Phi, double* Psi) {
  constexpr int CW = cw<CPLX>();
  const size_t n = in.n;
  Engine<CPLX> e(c, n, in.naux);
  double* dK = static_cast<double*>(wsbuf(c, WS_SG_KIN, std::max<size_t>(1, n * n) * sizeof(double)));
  RR_HIP(hipMemcpyAsync(dK, Kh.data(), n * n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  zero_dev(c, Phi, n * n * CW);
  zero_dev(c, Psi, n * n * CW);
  Toolkit t = build_toolkit<CPLX>(e, in.dB, dK);
  HostMat K{n, Kh.data()};
  SigmaRun<CPLX> run(e, in, t, dK, K);
  if (!hx) {
    run.jk_phi(Phi);
    run.jk_psi(Psi);
  } else {
    run.hx_phi(Phi);
    run.hx_psi(Psi);
    if (in.pair_of && !in.pair_of->empty()) {
      run.l12_phi(Phi);
      run.l12_psi(Psi);
    }
  }
}

using Pairs = std::vector<std::pair<size_t, size_t>>;

// w_I = 1/2 [(Phi - Phi^T) + (Psi - Psi^T)]_pq  (real kappa)
void sigma_real(Context& c, const SigmaInputs& in, bool hx, const Pairs& pairs, const double* v, double* w) {
  const size_t n = in.n;
  std::vector<double> K(n * n, 0.0);
  for (size_t i = 0; i < pairs.size(); ++i) {
    K[pairs[i].first * n + pairs[i].second] = v[i];
    K[pairs[i].second * n + pairs[i].first] = -v[i];
  }
  double* Phi = static_cast<double*>(wsbuf(c, WS_SG_PHI, std::max<size_t>(1, n * n) * sizeof(double)));
  double* Psi = static_cast<double*>(wsbuf(c, WS_SG_PSI, std::max<size_t>(1, n * n) * sizeof(double)));
  phi_psi<false>(c, in, hx, K, Phi, Psi);
  std::vector<double> phi(n * n), psi(n * n);
  RR_HIP(hipMemcpyAsync(phi.data(), Phi, n * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipMemcpyAsync(psi.data(), Psi, n * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  for (size_t i = 0; i < pairs.size(); ++i) {
    const size_t p = pairs[i].first, q = pairs[i].second;
    w[i] = 0.5 * ((phi[p * n + q] - phi[q * n + p]) + (psi[p * n + q] - psi[q * n + p]));
  }
}

// Joint [t; y] product (JkOnlySigmaHessianVectorRi.cpp:525-573).
void sigma_joint(Context& c, const SigmaInputs& in, bool hx, const Pairs& pairs, const double* v, double* w) {
  using Cx = std::complex<double>;
  const size_t n = in.n, np = pairs.size();
  std::vector<double> Kt(n * n, 0.0), Ky(n * n, 0.0);
  for (size_t i = 0; i < np; ++i) {
    const size_t p = pairs[i].first, q = pairs[i].second;
    Kt[p * n + q] = v[i];
    Kt[q * n + p] = -v[i];
    Ky[p * n + q] = v[np + i];
    Ky[q * n + p] = v[np + i];
  }
  double* Phi = static_cast<double*>(wsbuf(c, WS_SG_PHI, std::max<size_t>(1, n * n) * 2 * sizeof(double)));
  double* Psi = static_cast<double*>(wsbuf(c, WS_SG_PSI, std::max<size_t>(1, n * n) * 2 * sizeof(double)));
  std::vector<Cx> phit(n * n), psit(n * n), phiy(n * n), psiy(n * n);
  phi_psi<true>(c, in, hx, Kt, Phi, Psi);
  RR_HIP(hipMemcpyAsync(phit.data(), Phi, n * n * sizeof(Cx), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipMemcpyAsync(psit.data(), Psi, n * n * sizeof(Cx), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  phi_psi<true>(c, in, hx, Ky, Phi, Psi);
  RR_HIP(hipMemcpyAsync(phiy.data(), Phi, n * n * sizeof(Cx), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipMemcpyAsync(psiy.data(), Psi, n * n * sizeof(Cx), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  const Cx im(0.0, 1.0);
  for (size_t i = 0; i < np; ++i) {
    const size_t p = pairs[i].first, q = pairs[i].second;
    const size_t pq = p * n + q, qp = q * n + p;
    const double wt = (0.5 * ((phit[pq] - phit[qp]) + (psit[pq] - psit[qp]))).real() +
                      (0.5 * im * ((phiy[pq] - phiy[qp]) + (psiy[pq] - psiy[qp]))).real();
    const double wy = (0.5 * im * ((psit[pq] + psit[qp]) + (phit[pq] + phit[qp]))).real() -
                      (0.5 * ((phiy[pq] + phiy[qp]) + (psiy[pq] + psiy[qp]))).real();
    w[i] = wt;
    w[np + i] = wy;
  }
}

}  // namespace
}  // namespace rr

// ================================================================================================
// C ABI
// ================================================================================================
using rr::Context;
using rr::Error;
using rr::guarded;

namespace {

// Device copy of a caller array (host or device), `count` doubles.
const double* dev_in(Context& c, int id, const double* p, size_t count) { return rr::in_buf(c, id, p, count); }

rr::Pairs read_pairs(size_t n, size_t n_pairs, const size_t* pairs) {
  rr::Pairs out(n_pairs);
  for (size_t i = 0; i < n_pairs; ++i) {
    out[i] = {pairs[2 * i], pairs[2 * i + 1]};
    RR_REQUIRE(out[i].first < n && out[i].second < n, RR_ERR_ARG, "pair index out of range");
  }
  return out;
}

std::vector<size_t> read_pair_of(size_t n, const size_t* pair_of) {
  std::vector<size_t> pr;
  if (!pair_of) return pr;
  pr.assign(pair_of, pair_of + n);
  for (size_t i = 0; i < n; ++i) RR_REQUIRE(pr[i] < n, RR_ERR_ARG, "pair_of index out of range");
  return pr;
}

int sigma_entry(rr_handle h, bool cplx, bool hx, const char* who, size_t n, const double* h_mo, const double* b,
                size_t n_aux, const double* occupations, const double* two_rdm_h, const double* two_rdm_x,
                const double* fock, const size_t* pair_of, const double* two_rdm_l1, const double* two_rdm_l2,
                size_t n_pairs, const size_t* pairs, const double* v, double* w) {
  return guarded(h, [&](Context& c) {
    const int cwv = cplx ? 2 : 1;
    RR_REQUIRE(h_mo && b && occupations && two_rdm_h && two_rdm_x && w && (n_pairs == 0 || (pairs && v)), RR_ERR_ARG,
               std::string(who) + ": null argument");
    RR_REQUIRE(!hx || fock, RR_ERR_ARG, std::string(who) + ": null argument");
    RR_REQUIRE(!pair_of || (two_rdm_l1 && two_rdm_l2), RR_ERR_ARG, std::string(who) + ": null argument");
    rr::SigmaInputs in;
    in.n = n, in.naux = n_aux;
    in.dB = dev_in(c, rr::WS_SG_B, b, n_aux * n * n * cwv);
    in.dh = dev_in(c, rr::WS_SG_H, h_mo, n * n * cwv);
    in.docc = dev_in(c, rr::WS_SG_OCC, occupations, n);
    // couplings are needed on the host (weight tables) and on the device (Uh/V products)
    std::vector<double> hc(n * n), xc(n * n), l1, l2;
    RR_HIP(hipMemcpy(hc.data(), two_rdm_h, n * n * sizeof(double), hipMemcpyDefault));
    RR_HIP(hipMemcpy(xc.data(), two_rdm_x, n * n * sizeof(double), hipMemcpyDefault));
    in.hc = {n, hc.data()};
    in.xc = {n, xc.data()};
    in.dhc = dev_in(c, rr::WS_SG_HC, two_rdm_h, n * n);
    in.dxc = dev_in(c, rr::WS_SG_XC, two_rdm_x, n * n);
    if (hx) in.dfock = dev_in(c, rr::WS_SG_FOCKIN, fock, n * n * cwv);
    std::vector<size_t> pr = read_pair_of(n, pair_of);
    if (pair_of) {
      l1.resize(n * n), l2.resize(n * n);
      RR_HIP(hipMemcpy(l1.data(), two_rdm_l1, n * n * sizeof(double), hipMemcpyDefault));
      RR_HIP(hipMemcpy(l2.data(), two_rdm_l2, n * n * sizeof(double), hipMemcpyDefault));
      in.pair_of = &pr;
      in.l1 = {n, l1.data()};
      in.l2 = {n, l2.data()};
    }
    rr::Pairs pp = read_pairs(n, n_pairs, pairs);
    std::vector<double> vh(n_pairs * cwv), wh(n_pairs * cwv);
    if (n_pairs) RR_HIP(hipMemcpy(vh.data(), v, vh.size() * sizeof(double), hipMemcpyDefault));
    if (cplx)
      rr::sigma_joint(c, in, hx, pp, vh.data(), wh.data());
    else
      rr::sigma_real(c, in, hx, pp, vh.data(), wh.data());
    if (n_pairs) RR_HIP(hipMemcpy(w, wh.data(), wh.size() * sizeof(double), hipMemcpyDefault));
  });
}

}  // namespace

extern "C" {

int rr_ri_hx_fock_gemm(rr_handle h, int is_complex, size_t n, size_t n_aux, const double* b, const double* h_mo,
                       const double* occupations, const double* two_rdm_h, const double* two_rdm_x,
                       const size_t* pair_of, const double* two_rdm_l1, const double* two_rdm_l2, double* fock) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b && h_mo && occupations && two_rdm_h && two_rdm_x && fock, RR_ERR_ARG, "riHxFockGemm: null argument");
    RR_REQUIRE(!pair_of || (two_rdm_l1 && two_rdm_l2), RR_ERR_ARG,
               "riHxFockGemm: pair_of/two_rdm_l1/two_rdm_l2 dimensions inconsistent with h");
    const int cwv = is_complex ? 2 : 1;
    const double* dB = dev_in(c, rr::WS_SG_B, b, n_aux * n * n * cwv);
    const double* dh = dev_in(c, rr::WS_SG_H, h_mo, n * n * cwv);
    const double* docc = dev_in(c, rr::WS_SG_OCC, occupations, n);
    const double* dh2 = dev_in(c, rr::WS_SG_HC, two_rdm_h, n * n);
    const double* dx2 = dev_in(c, rr::WS_SG_XC, two_rdm_x, n * n);
    const int64_t* dpair = nullptr;
    const double* dcmat = nullptr;
    if (pair_of) {
      std::vector<size_t> pr = read_pair_of(n, pair_of);
      std::vector<double> l1(n * n), l2(n * n), cm(n * n);
      RR_HIP(hipMemcpy(l1.data(), two_rdm_l1, n * n * sizeof(double), hipMemcpyDefault));
      RR_HIP(hipMemcpy(l2.data(), two_rdm_l2, n * n * sizeof(double), hipMemcpyDefault));
      // c(s,q) = l1(s, qbar) + l2(s, q)
      for (size_t s = 0; s < n; ++s)
        for (size_t q = 0; q < n; ++q) cm[s * n + q] = l1[s * n + pr[q]] + l2[s * n + q];
      std::vector<int64_t> pi(pr.begin(), pr.end());
      dpair = rr::upload_tab(c, rr::WS_SG_PAIR, pi);
      dcmat = rr::upload_tab(c, rr::WS_SG_CMAT, cm);
    }
    rr::OutBuf of = rr::out_buf(c, rr::WS_SG_FOCKOUT, fock, n * n * cwv);
    if (is_complex)
      rr::hx_fock<true>(c, n, n_aux, dB, dh, docc, dh2, dx2, dpair, dcmat, of.dev);
    else
      rr::hx_fock<false>(c, n, n_aux, dB, dh, docc, dh2, dx2, dpair, dcmat, of.dev);
    rr::finish_out(c, of);
  });
}

int rr_ri_sigma_toolkit(rr_handle h, int is_complex, size_t n, size_t n_aux, const double* b, const double* kappa,
                        double* kb1, double* kb2, double* bdiag, double* kb1diag, double* kb2diag) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b && kappa, RR_ERR_ARG, "buildRiSigmaHessianToolkit: null argument");
    const int cwv = is_complex ? 2 : 1;
    const double* dB = dev_in(c, rr::WS_SG_B, b, n_aux * n * n * cwv);
    const double* dK = dev_in(c, rr::WS_SG_KIN, kappa, n * n);
    auto run = [&](auto tag) {
      constexpr bool CPLX = decltype(tag)::value;
      rr::Engine<CPLX> e(c, n, n_aux);
      rr::Toolkit t = rr::build_toolkit<CPLX>(e, dB, dK);
      const size_t cnt = n_aux * n * n * cwv;
      if (kb1 && cnt) RR_HIP(hipMemcpyAsync(kb1, t.kb1, cnt * sizeof(double), hipMemcpyDefault, c.stream));
      if (kb2 && cnt) RR_HIP(hipMemcpyAsync(kb2, t.kb2, cnt * sizeof(double), hipMemcpyDefault, c.stream));
      // diagonals: padded rows of length ldx on device -> (n_aux, n) for the caller
      const size_t spitch = e.ldx * cwv * sizeof(double), dpitch = n * cwv * sizeof(double);
      double* outs[3] = {bdiag, kb1diag, kb2diag};
      double* srcs[3] = {t.bdiag, t.kb1diag, t.kb2diag};
      for (int i = 0; i < 3; ++i)
        if (outs[i] && n && n_aux)
          RR_HIP(hipMemcpy2DAsync(outs[i], dpitch, srcs[i], spitch, dpitch, n_aux, hipMemcpyDefault, c.stream));
    };
    if (is_complex)
      run(std::true_type{});
    else
      run(std::false_type{});
  });
}

int rr_ri_pair_contract(rr_handle h, int is_complex, size_t n, size_t n_aux, double sign, const rr_pair_op* f,
                        const rr_pair_op* g, double* target) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(f && g && target, RR_ERR_ARG, "pairContract: null argument");
    const int cwv = is_complex ? 2 : 1;
    auto to_op = [&](const rr_pair_op* o, int base) {
      rr::Op op;
      if (o->s) op.s = dev_in(c, base, o->s, n_aux * n * n * cwv);
      if (o->s) {
        RR_REQUIRE(o->idx, RR_ERR_ARG, "pairContract: an operand with a source tensor needs an index table");
        op.mode = 2;
        std::vector<uint32_t> idx(n * n);
        RR_HIP(hipMemcpy(idx.data(), o->idx, n * n * sizeof(uint32_t), hipMemcpyDefault));
        for (uint32_t v : idx) RR_REQUIRE(v < n * n, RR_ERR_ARG, "pairContract: index table entry out of range");
        op.idx = rr::upload_tab(c, base + 1, idx);
      }
      if (o->w) op.w = dev_in(c, base + 2, o->w, n * n);
      if (o->row) op.row = dev_in(c, base + 3, o->row, n_aux * n * cwv);
      if (o->post) op.post = dev_in(c, base + 4, o->post, n * n);
      return op;
    };
    rr::Op fo = to_op(f, rr::WS_SG_OPF), go = to_op(g, rr::WS_SG_OPG);
    double* dt = const_cast<double*>(dev_in(c, rr::WS_SG_TARGET2, target, n * n * cwv));
    if (is_complex) {
      rr::Engine<true> e(c, n, n_aux);
      e.pair_contract(dt, sign, fo, go);
    } else {
      rr::Engine<false> e(c, n, n_aux);
      e.pair_contract(dt, sign, fo, go);
    }
    if (dt != target) RR_HIP(hipMemcpyAsync(target, dt, n * n * cwv * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  });
}

int rr_ri_diag_contract(rr_handle h, int is_complex, size_t n, size_t n_aux, const double* s, const double* ua,
                        const double* ub, double* target) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(s && target, RR_ERR_ARG, "diagContract: null argument");
    const int cwv = is_complex ? 2 : 1;
    const double* dS = dev_in(c, rr::WS_SG_B, s, n_aux * n * n * cwv);
    const double* dUa = ua ? dev_in(c, rr::WS_SG_U1, ua, n_aux * n * cwv) : nullptr;
    const double* dUb = ub ? dev_in(c, rr::WS_SG_U2, ub, n_aux * n * cwv) : nullptr;
    double* dt = const_cast<double*>(dev_in(c, rr::WS_SG_TARGET2, target, n * n * cwv));
    if (is_complex) {
      rr::Engine<true> e(c, n, n_aux);
      e.diag_contract(dt, dS, dUa, dUb);
    } else {
      rr::Engine<false> e(c, n, n_aux);
      e.diag_contract(dt, dS, dUa, dUb);
    }
    if (dt != target) RR_HIP(hipMemcpyAsync(target, dt, n * n * cwv * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  });
}

int rr_ri_accum_u(rr_handle h, int is_complex, size_t n, size_t n_aux, double alpha, const double* d, const double* wm,
                  double* u) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(d && wm && u, RR_ERR_ARG, "accumU: null argument");
    const int cwv = is_complex ? 2 : 1;
    const size_t ld = rr::roundup(std::max<size_t>(n, 1), 8);
    const double* dD = dev_in(c, rr::WS_SG_B, d, n_aux * n * cwv);
    double* dpad = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_DPAD, std::max<size_t>(1, n_aux * ld) * cwv * sizeof(double)));
    const double* dW = dev_in(c, rr::WS_SG_HC, wm, n * n);
    double* du = const_cast<double*>(dev_in(c, rr::WS_SG_U1, u, n_aux * n * cwv));
    if (is_complex) {
      hipLaunchKernelGGL((rr::k_pad<true>), dim3(rr::grid_for(n_aux * ld)), dim3(256), 0, c.stream, dD, 2, n, 0, n_aux, n,
                         ld, n_aux, dpad);
      rr::Engine<true> e(c, n, n_aux);
      e.accum_u(du, alpha, dpad, dW);
    } else {
      hipLaunchKernelGGL((rr::k_pad<false>), dim3(rr::grid_for(n_aux * ld)), dim3(256), 0, c.stream, dD, 1, n, 0, n_aux, n,
                         ld, n_aux, dpad);
      rr::Engine<false> e(c, n, n_aux);
      e.accum_u(du, alpha, dpad, dW);
    }
    if (du != u) RR_HIP(hipMemcpyAsync(u, du, n_aux * n * cwv * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  });
}

int rr_jk_only_sigma_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b, size_t n_aux,
                                       const double* occupations, const double* two_rdm_h, const double* two_rdm_x,
                                       size_t n_pairs, const size_t* pairs, const double* v, double* w) {
  return sigma_entry(h, false, false, "jkOnlySigmaHessianVectorRi", n, h_mo, b, n_aux, occupations, two_rdm_h, two_rdm_x,
                     nullptr, nullptr, nullptr, nullptr, n_pairs, pairs, v, w);
}

int rr_jk_only_sigma_joint_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b, size_t n_aux,
                                             const double* occupations, const double* two_rdm_h,
                                             const double* two_rdm_x, size_t n_pairs, const size_t* pairs,
                                             const double* v, double* w) {
  return sigma_entry(h, true, false, "jkOnlySigmaJointHessianVectorRi", n, h_mo, b, n_aux, occupations, two_rdm_h,
                     two_rdm_x, nullptr, nullptr, nullptr, nullptr, n_pairs, pairs, v, w);
}

int rr_hartree_exchange_sigma_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b,
                                                size_t n_aux, const double* occupations, const double* two_rdm_h,
                                                const double* two_rdm_x, const double* fock, const size_t* pair_of,
                                                const double* two_rdm_l1, const double* two_rdm_l2, size_t n_pairs,
                                                const size_t* pairs, const double* v, double* w) {
  return sigma_entry(h, false, true, "hartreeExchangeSigmaHessianVectorRi", n, h_mo, b, n_aux, occupations, two_rdm_h,
                     two_rdm_x, fock, pair_of, two_rdm_l1, two_rdm_l2, n_pairs, pairs, v, w);
}

int rr_hartree_exchange_sigma_joint_hessian_vector_ri(rr_handle h, size_t n, const double* h_mo, const double* b,
                                                      size_t n_aux, const double* occupations, const double* two_rdm_h,
                                                      const double* two_rdm_x, const double* fock,
                                                      const size_t* pair_of, const double* two_rdm_l1,
                                                      const double* two_rdm_l2, size_t n_pairs, const size_t* pairs,
                                                      const double* v, double* w) {
  return sigma_entry(h, true, true, "hartreeExchangeSigmaJointHessianVectorRi", n, h_mo, b, n_aux, occupations,
                     two_rdm_h, two_rdm_x, fock, pair_of, two_rdm_l1, two_rdm_l2, n_pairs, pairs, v, w);
}

// riJkOnlyFockMatrix / riPnofFockMatrix / riPccdFockMatrix body (RiJkOnlyFock.cpp:14-33,
// RiPnofFock.cpp:14-30, RiPccdFock.cpp:14-31): C_ukb = V C, B^MO = buildRiMoEri(C_ukb) kept on the
// device, h_mo = C^H h C, then the generalized Fock GEMM; the functional enters only through the
// couplings (and pair terms) the caller passes.
int rr_ri_mo_fock_rkb(rr_handle h, int slot_ll, int slot_ss, size_t nl, size_t ns, const double* h_rkb,
                      const double* v_total, size_t n_rkb, const double* c_rkb, size_t n_mo, const double* occupations,
                      const double* two_rdm_h, const double* two_rdm_x, const size_t* pair_of, const double* two_rdm_l1,
                      const double* two_rdm_l2, double* fock) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(h_rkb && v_total && c_rkb && occupations && two_rdm_h && two_rdm_x && fock, RR_ERR_ARG,
               "riJkOnlyFockMatrix: null argument");
    RR_REQUIRE(!pair_of || (two_rdm_l1 && two_rdm_l2), RR_ERR_ARG,
               "riHxFockGemm: pair_of/two_rdm_l1/two_rdm_l2 dimensions inconsistent with h");
    const size_t nu = 2 * nl + 2 * ns;
    const size_t naux = rr::slot_naux(c, slot_ll);
    const double* dh = dev_in(c, rr::WS_SG_H2IN, h_rkb, 2 * n_rkb * n_rkb);
    const double* dv = dev_in(c, rr::WS_SG_V2, v_total, 2 * nu * n_rkb);
    const double* dc = dev_in(c, rr::WS_SG_C2, c_rkb, 2 * n_rkb * n_mo);
    double* cukb = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_CUKB, std::max<size_t>(1, 2 * nu * n_mo) * sizeof(double)));
    rr::dgemm_small(c, true, 'N', 'N', nu, n_mo, n_rkb, dv, n_rkb, dc, n_mo, cukb, n_mo);  // C_ukb = V C
    double* mo = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_MO, std::max<size_t>(1, 2 * naux * n_mo * n_mo) * sizeof(double)));
    rr::mo_tensor_rkb(c, slot_ll, slot_ss, nl, ns, cukb, n_mo, mo);
    double* tmp = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_HTMP, std::max<size_t>(1, 2 * n_rkb * n_mo) * sizeof(double)));
    double* hmo = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_HMO, std::max<size_t>(1, 2 * n_mo * n_mo) * sizeof(double)));
    rr::dgemm_small(c, true, 'N', 'N', n_rkb, n_mo, n_rkb, dh, n_rkb, dc, n_mo, tmp, n_mo);     // h C
    rr::dgemm_small(c, true, 'C', 'N', n_mo, n_mo, n_rkb, dc, n_mo, tmp, n_mo, hmo, n_mo);      // C^H h C
    const double* docc = dev_in(c, rr::WS_SG_OCC, occupations, n_mo);
    const double* dh2 = dev_in(c, rr::WS_SG_HC, two_rdm_h, n_mo * n_mo);
    const double* dx2 = dev_in(c, rr::WS_SG_XC, two_rdm_x, n_mo * n_mo);
    const int64_t* dpair = nullptr;
    const double* dcmat = nullptr;
    if (pair_of) {
      std::vector<size_t> pr = read_pair_of(n_mo, pair_of);
      std::vector<double> l1(n_mo * n_mo), l2(n_mo * n_mo), cm(n_mo * n_mo);
      RR_HIP(hipMemcpy(l1.data(), two_rdm_l1, l1.size() * sizeof(double), hipMemcpyDefault));
      RR_HIP(hipMemcpy(l2.data(), two_rdm_l2, l2.size() * sizeof(double), hipMemcpyDefault));
      for (size_t s = 0; s < n_mo; ++s)
        for (size_t q = 0; q < n_mo; ++q) cm[s * n_mo + q] = l1[s * n_mo + pr[q]] + l2[s * n_mo + q];
      std::vector<int64_t> pi(pr.begin(), pr.end());
      dpair = rr::upload_tab(c, rr::WS_SG_PAIR, pi);
      dcmat = rr::upload_tab(c, rr::WS_SG_CMAT, cm);
    }
    rr::OutBuf of = rr::out_buf(c, rr::WS_SG_FOCKOUT, fock, 2 * n_mo * n_mo);
    rr::hx_fock<true>(c, n_mo, naux, mo, hmo, docc, dh2, dx2, dpair, dcmat, of.dev);
    rr::finish_out(c, of);
  });
}

// riJkOnlyFockMatrixNonRel / riPnofFockMatrixNonRel / riPccdFockMatrixNonRel body
// (RiNonRelFullOptFock.cpp:53-98): spatial half of c_current (n x 2n), embedded spin-orbital MO
// tensor, spin-blocked h_mo = C^T h C (buildSpinOneElectron, :38-49), then the generalized Fock GEMM.
int rr_ri_mo_fock_nonrel(rr_handle h, int slot, size_t n_spatial, const double* h_core_ao, const double* c_current,
                         const double* occupations, const double* two_rdm_h, const double* two_rdm_x,
                         const size_t* pair_of, const double* two_rdm_l1, const double* two_rdm_l2, double* fock) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(h_core_ao && c_current && occupations && two_rdm_h && two_rdm_x && fock, RR_ERR_ARG,
               "riJkOnlyFockMatrixNonRel: null argument");
    RR_REQUIRE(!pair_of || (two_rdm_l1 && two_rdm_l2), RR_ERR_ARG,
               "riHxFockGemm: pair_of/two_rdm_l1/two_rdm_l2 dimensions inconsistent with h");
    const size_t ns = n_spatial, nt = 2 * ns;
    const size_t naux = rr::slot_naux(c, slot);
    const double* dh = dev_in(c, rr::WS_SG_H2IN, h_core_ao, ns * ns);
    const double* dc = dev_in(c, rr::WS_SG_C2, c_current, ns * nt);
    double* csp = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_CSP, std::max<size_t>(1, ns * ns) * sizeof(double)));
    hipLaunchKernelGGL((rr::k_pad<false>), dim3(rr::grid_for(ns * ns)), dim3(256), 0, c.stream, dc, 1, nt, 0, ns, ns, ns,
                       ns, csp);  // spatialHalf: first n_spatial columns
    double* mo = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_MO, std::max<size_t>(1, naux * nt * nt) * sizeof(double)));
    rr::mo_tensor_nonrel_embedded(c, slot, ns, csp, mo);
    double* tmp = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_HTMP, std::max<size_t>(1, nt * ns) * sizeof(double)));
    double* hmo = static_cast<double*>(rr::wsbuf(c, rr::WS_SG_HMO, std::max<size_t>(1, nt * nt) * sizeof(double)));
    rr::dgemm_small(c, false, 'T', 'N', nt, ns, ns, dc, nt, dh, ns, tmp, ns);  // C^T h
    rr::dgemm_small(c, false, 'N', 'N', nt, nt, ns, tmp, ns, dc, nt, hmo, nt);  // C^T h C
    hipLaunchKernelGGL(rr::k_spin_blocks, dim3(rr::grid_for(nt * nt)), dim3(256), 0, c.stream, hmo, ns);
    RR_HIP(hipGetLastError());
    const double* docc = dev_in(c, rr::WS_SG_OCC, occupations, nt);
    const double* dh2 = dev_in(c, rr::WS_SG_HC, two_rdm_h, nt * nt);
    const double* dx2 = dev_in(c, rr::WS_SG_XC, two_rdm_x, nt * nt);
    const int64_t* dpair = nullptr;
    const double* dcmat = nullptr;
    if (pair_of) {
      std::vector<size_t> pr = read_pair_of(nt, pair_of);
      std::vector<double> l1(nt * nt), l2(nt * nt), cm(nt * nt);
      RR_HIP(hipMemcpy(l1.data(), two_rdm_l1, l1.size() * sizeof(double), hipMemcpyDefault));
      RR_HIP(hipMemcpy(l2.data(), two_rdm_l2, l2.size() * sizeof(double), hipMemcpyDefault));
      for (size_t s = 0; s < nt; ++s)
        for (size_t q = 0; q < nt; ++q) cm[s * nt + q] = l1[s * nt + pr[q]] + l2[s * nt + q];
      std::vector<int64_t> pi(pr.begin(), pr.end());
      dpair = rr::upload_tab(c, rr::WS_SG_PAIR, pi);
      dcmat = rr::upload_tab(c, rr::WS_SG_CMAT, cm);
    }
    rr::OutBuf of = rr::out_buf(c, rr::WS_SG_FOCKOUT, fock, nt * nt);
    rr::hx_fock<false>(c, nt, naux, mo, hmo, docc, dh2, dx2, dpair, dcmat, of.dev);
    rr::finish_out(c, of);
  });
}

}  // extern "C"
