// librerdmft_gpu, 4c RKB Cholesky path: the Fock build over RKB Cholesky vectors (Coulomb with the
// cross-flavor small-component terms, exchange K_fg = sum_L B^f_L P_fg B^g_L per flavor block) and
// the transform of the vectors to the (positive-energy) MO basis with the optional recompression.
// FP64 complex throughout.
//
// Layout: a vector set (RkbChol in rr_core.hpp) keeps the five components of every vector L padded
// to rows of ld = roundup(nl, 8): nat[comp][L][a][m] as given, tr[comp][L][b][s] = B[comp][L][s][b]
// for the three flavor components. All products are NT GEMMs of rr_gemm.hpp:
//   exchange, per flavor g and chunk of vectors:
//     Tt_L[b][(f,m)] = sum_s Bt^g_L[b][s] P_fg[m][s]                    (batched over L)
//     K_fg^T[b][a]  += sum_L sum_m Tt_L[b][(f,m)] B^f_L[a][m]           (k-batched over L, split-K)
//   MO transform, per (bra, ket, component) term and chunk of vectors:
//     X_L[q][a] = sum_m C_ket^T[q][m] B_L[a][m],   W_L^T[q][p] += sum_a C_bra^H[p][a] X_L[q][a]
#include "rr_core.hpp"

namespace rr {
namespace {

enum RkbWs {
  WS_RK_STAGE = WS_SIGMA_BASE + 3000, WS_RK_MT, WS_RK_T, WS_RK_JPART, WS_RK_J, WS_RK_PST, WS_RK_TT, WS_RK_KPART,
  WS_RK_K, WS_RK_P, WS_RK_H, WS_RK_OUT, WS_RK_C, WS_RK_CT, WS_RK_CH, WS_RK_X, WS_RK_A, WS_RK_G, WS_RK_W, WS_RK_U,
  WS_RK_AT, WS_RK_AP, WS_RK_S, WS_RK_XF, WS_RK_CT2, WS_RK_OCC, WS_RK_FP, WS_RK_WF, WS_RK_HERM
};

#define RR_GRID_LOOP(i, total) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (total); i += (size_t)gridDim.x * blockDim.x)

constexpr int kLarge = 0, kSmall0 = 1, kSmall1 = 2, kCrossF = 3, kCrossB = 4;
// vector component carried by flavor f (La, Lb share the large vectors)
inline int flavor_comp(int f) { return f < 2 ? kLarge : f - 1; }

// Stores Lc vectors of one component: nat[L][r][c] = src[L][r*nl + c], tr[L][r][c] = src[L][c*nl + r]
// (complex; a real source is promoted), zero for c >= nl.
__global__ void k_rkb_store(const double* src, int real_src, size_t lc, size_t nl, size_t ld, double* nat, double* tr) {
  const size_t per = nl * ld;
  RR_GRID_LOOP(i, lc * per) {
    const size_t L = i / per, r = (i % per) / ld, c = i % ld;
    double nr = 0.0, ni = 0.0, tr_ = 0.0, ti = 0.0;
    if (c < nl) {
      const size_t en = L * nl * nl + r * nl + c, et = L * nl * nl + c * nl + r;
      if (real_src) {
        nr = src[en], tr_ = src[et];
      } else {
        nr = src[2 * en], ni = src[2 * en + 1], tr_ = src[2 * et], ti = src[2 * et + 1];
      }
    }
    nat[2 * i] = nr, nat[2 * i + 1] = ni;
    if (tr) tr[2 * i] = tr_, tr[2 * i + 1] = ti;
  }
}

// Transposed density blocks for the Coulomb traces: Mt[comp][q][s] (ld-padded) =
//   large: P00[s][q] + P11[s][q]; small0: P22[s][q]; small1: P33[s][q]; cross_f: P32[s][q]; cross_b: P23[s][q]
__global__ void k_rkb_mt(const double* P, size_t nl, size_t ld, double* mt) {
  const size_t N = 4 * nl, per = nl * ld;
  RR_GRID_LOOP(i, 5 * per) {
    const int comp = static_cast<int>(i / per);
    const size_t q = (i % per) / ld, s = i % ld;
    double re = 0.0, im = 0.0;
    if (s < nl) {
      auto add = [&](int bf, int bg) {
        const size_t e = (bf * nl + s) * N + bg * nl + q;
        re += P[2 * e], im += P[2 * e + 1];
      };
      if (comp == kLarge) {
        add(0, 0);
        add(1, 1);
      } else if (comp == kSmall0) {
        add(2, 2);
      } else if (comp == kSmall1) {
        add(3, 3);
      } else if (comp == kCrossF) {
        add(3, 2);
      } else {
        add(2, 3);
      }
    }
    mt[2 * i] = re, mt[2 * i + 1] = im;
  }
}

// t_L = sum_comp sum_{q,s} nat[comp][L][q][s] * Mt[comp][q][s]   (one workgroup per vector)
__global__ __launch_bounds__(256) void k_rkb_trace(const double* nat, size_t nchol, size_t per, const double* mt,
                                                   double* t) {
  __shared__ double sr[256], si[256];
  for (size_t L = blockIdx.x; L < nchol; L += gridDim.x) {
    double ar = 0.0, ai = 0.0;
    for (int comp = 0; comp < 5; ++comp) {
      const double* v = nat + 2 * ((size_t)comp * nchol + L) * per;
      const double* m = mt + 2 * (size_t)comp * per;
      for (size_t e = threadIdx.x; e < per; e += blockDim.x) {
        const double vr = v[2 * e], vi = v[2 * e + 1], mr = m[2 * e], mi = m[2 * e + 1];
        ar += vr * mr - vi * mi;
        ai += vr * mi + vi * mr;
      }
    }
    sr[threadIdx.x] = ar, si[threadIdx.x] = ai;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sr[threadIdx.x] += sr[threadIdx.x + s], si[threadIdx.x] += si[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) t[2 * L] = sr[0], t[2 * L + 1] = si[0];
    __syncthreads();
  }
}

// part[z][comp][e] = sum_{L in split z} t_L nat[comp][L][e]
__global__ void k_rkb_j_part(const double* nat, size_t nchol, size_t per, const double* t, size_t lper, double* part) {
  const size_t z = blockIdx.y;
  const size_t l0 = z * lper, l1 = min(nchol, l0 + lper);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < 5 * per; i += (size_t)gridDim.x * blockDim.x) {
    const size_t comp = i / per, e = i % per;
    double ar = 0.0, ai = 0.0;
    for (size_t L = l0; L < l1; ++L) {
      const double* v = nat + 2 * ((comp * nchol + L) * per + e);
      const double tr_ = t[2 * L], ti = t[2 * L + 1];
      ar += tr_ * v[0] - ti * v[1];
      ai += tr_ * v[1] + ti * v[0];
    }
    part[2 * (z * 5 * per + i)] = ar;
    part[2 * (z * 5 * per + i) + 1] = ai;
  }
}
__global__ void k_rkb_j_reduce(const double* part, int nsplit, size_t count, double* j) {
  RR_GRID_LOOP(i, 2 * count) {
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += part[z * 2 * count + i];
    j[i] = s;
  }
}

// Pstack_g[(f*ld + m)][s] = P[f*nl + m][g*nl + s] for m, s < nl (zero padding), 4*ld rows of ld.
__global__ void k_rkb_pstack(const double* P, size_t nl, size_t ld, int g, double* ps) {
  const size_t N = 4 * nl;
  RR_GRID_LOOP(i, 4 * ld * ld) {
    const size_t row = i / ld, s = i % ld, f = row / ld, m = row % ld;
    double re = 0.0, im = 0.0;
    if (m < nl && s < nl) {
      const size_t e = (f * nl + m) * N + g * nl + s;
      re = P[2 * e], im = P[2 * e + 1];
    }
    ps[2 * i] = re, ps[2 * i + 1] = im;
  }
}

// K[(f*nl + a)*N + g*nl + b] += sum_z part[z][b][a]
__global__ void k_rkb_kadd(const double* part, int nsplit, size_t nl, int f, int g, double* K) {
  const size_t N = 4 * nl, nn = nl * nl;
  RR_GRID_LOOP(i, nn) {
    const size_t b = i / nl, a = i % nl;
    double sr = 0.0, si = 0.0;
    for (int z = 0; z < nsplit; ++z) sr += part[2 * (z * nn + i)], si += part[2 * (z * nn + i) + 1];
    const size_t e = (f * nl + a) * N + g * nl + b;
    K[2 * e] += sr, K[2 * e + 1] += si;
  }
}

// F = h + J - K with J_ff = J[comp(f)], F(Sa,Sb) += J_cross_f, F(Sb,Sa) += J_cross_b
__global__ void k_rkb_assemble(const double* h, const double* J, const double* K, size_t nl, size_t ld, double* F) {
  const size_t N = 4 * nl, per = nl * ld;
  RR_GRID_LOOP(i, N * N) {
    const size_t r = i / N, c = i % N, f = r / nl, g = c / nl, a = r % nl, b = c % nl;
    double re = h[2 * i] - K[2 * i], im = h[2 * i + 1] - K[2 * i + 1];
    int comp = -1;
    if (f == g) comp = f < 2 ? kLarge : static_cast<int>(f) - 1;
    if (f == 2 && g == 3) comp = kCrossF;
    if (f == 3 && g == 2) comp = kCrossB;
    if (comp >= 0) {
      const size_t e = comp * per + a * ld + b;
      re += J[2 * e], im += J[2 * e + 1];
    }
    F[2 * i] = re, F[2 * i + 1] = im;
  }
}

// MO operands for flavor f: ct[q][m] = C[f*nl + m][nn + q], ch[p][a] = conj(C[f*nl + a][nn + p]),
// both m x ld (zero-padded), C is (4 nl) x nmo complex.
__global__ void k_rkb_cpack(const double* C, size_t nl, size_t nmo, size_t nn, size_t m, size_t ld, int f, double* ct,
                            double* ch) {
  RR_GRID_LOOP(i, m * ld) {
    const size_t q = i / ld, k = i % ld;
    double re = 0.0, im = 0.0;
    if (k < nl) {
      const size_t e = (f * nl + k) * nmo + nn + q;
      re = C[2 * e], im = C[2 * e + 1];
    }
    ct[2 * i] = re, ct[2 * i + 1] = im;
    ch[2 * i] = re, ch[2 * i + 1] = -im;
  }
}

__global__ void k_zero3(double* p, size_t count) {
  RR_GRID_LOOP(i, count) p[i] = 0.0;
}

// out[k][(nn + x)*nmo + nn + y] = ap[k][x*m + y] (zero elsewhere); ap rows of length lda.
__global__ void k_rkb_embed(const double* ap, size_t nk, size_t m, size_t lda, size_t nn, size_t nmo, double* out) {
  const size_t per = nmo * nmo;
  RR_GRID_LOOP(i, nk * per) {
    const size_t k = i / per, r = (i % per) / nmo, c = i % nmo;
    double re = 0.0, im = 0.0;
    if (r >= nn && c >= nn) {
      const size_t e = k * lda + (r - nn) * m + (c - nn);
      re = ap[2 * e], im = ap[2 * e + 1];
    }
    out[2 * i] = re, out[2 * i + 1] = im;
  }
}

// u[k][L] (row length ldl, zero-padded) = eig[first + k][L] (eigenvector rows)
__global__ void k_rkb_urows(const double* eig, size_t nchol, size_t first, size_t nk, size_t ldl, double* u) {
  RR_GRID_LOOP(i, nk * ldl) {
    const size_t k = i / ldl, L = i % ldl;
    double re = 0.0, im = 0.0;
    if (L < nchol) re = eig[2 * ((first + k) * nchol + L)], im = eig[2 * ((first + k) * nchol + L) + 1];
    u[2 * i] = re, u[2 * i + 1] = im;
  }
}
// at[j][L] (row length ldl) = a[L][j] (complex), j < cols
__global__ void k_rkb_transpose(const double* a, size_t rows, size_t lda, size_t cols, size_t ldl, double* at) {
  RR_GRID_LOOP(i, cols * ldl) {
    const size_t j = i / ldl, L = i % ldl;
    double re = 0.0, im = 0.0;
    if (L < rows) re = a[2 * (L * lda + j)], im = a[2 * (L * lda + j) + 1];
    at[2 * i] = re, at[2 * i + 1] = im;
  }
}

// Hermiticity of the large and the two small components (tr holds them transposed): per workgroup,
// out[2b] = max |nat - conj(tr)| and out[2b + 1] = max |nat| over its elements.
__global__ __launch_bounds__(256) void k_rkb_herm(const double* nat, const double* tr, size_t count, double* out) {
  __shared__ double sd[256], sm[256];
  double d = 0.0, m = 0.0;
  RR_GRID_LOOP(i, count) {
    const double ar = nat[2 * i], ai = nat[2 * i + 1], br = tr[2 * i], bi = tr[2 * i + 1];
    d = fmax(d, fmax(fabs(ar - br), fabs(ai + bi)));
    m = fmax(m, fmax(fabs(ar), fabs(ai)));
  }
  sd[threadIdx.x] = d, sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + s]);
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[2 * blockIdx.x] = sd[0], out[2 * blockIdx.x + 1] = sm[0];
}

// Half-transform operand of the occupied factor F = C sqrt(occ), per flavor f:
// fp[f][i][m] = C(f nl + m, i) sqrt(occ_i) from ct = C^T (n_occ x ldct complex); zero for i >= n_occ or
// m >= nl (4 rp rows of ld).
__global__ void k_rkb_fpack(const double* ct, size_t ldct, const double* occ, size_t n_occ, size_t nl, size_t rp,
                            size_t ld, double* fp) {
  RR_GRID_LOOP(i, 4 * rp * ld) {
    const size_t row = i / ld, m = i % ld, f = row / rp, o = row % rp;
    double re = 0.0, im = 0.0;
    if (o < n_occ && m < nl) {
      const double so = sqrt(occ[o]);
      const size_t e = o * ldct + f * nl + m;
      re = ct[2 * e] * so, im = ct[2 * e + 1] * so;
    }
    fp[2 * i] = re, fp[2 * i + 1] = im;
  }
}

// K (N x N complex) from the lower-tile partial sums of W W^H: K[r][c] = sum_z part[z][r][c] for r >= c
// and the conjugate of the mirrored element above the diagonal (the diagonal of a Hermitian K is real).
__global__ void k_rkb_kherm(const double* part, int nparts, size_t N, double* K) {
  const size_t nn = N * N;
  RR_GRID_LOOP(i, nn) {
    const size_t r = i / N, c = i % N;
    const size_t e = r >= c ? i : c * N + r;
    double sr = 0.0, si = 0.0;
    for (int z = 0; z < nparts; ++z) sr += part[2 * (z * nn + e)], si += part[2 * (z * nn + e) + 1];
    K[2 * i] = sr, K[2 * i + 1] = r > c ? si : (r == c ? 0.0 : -si);
  }
}

RkbChol& get_set(Context& c, int set) {
  RR_REQUIRE(set >= 0 && set < kRkbSets, RR_ERR_ARG, "RKB Cholesky set out of range");
  RkbChol& r = c.rkb[set];
  RR_REQUIRE(r.nat != nullptr, RR_ERR_STATE, "RKB Cholesky set is empty");
  return r;
}

// Split-K partial count for an (m x n) output with total flat k.
int splits_for(size_t m, size_t n, size_t ktot) {
  const size_t tiles = ((m + 63) / 64) * ((n + 63) / 64);
  size_t ns = std::max<size_t>(1, 1024 / std::max<size_t>(1, tiles));
  ns = std::min(ns, std::max<size_t>(1, ktot / 64));
  return (int)ns;
}

// J of every component from the density (k_rkb_mt, traces, split partial sums over the vectors):
// J[comp][a][b] (ld-padded rows, complex) as the reference accumulates t_L B_L (RkbCholesky.cpp:380-406).
double* rkb_coulomb(Context& c, const RkbChol& r, const double* dp) {
  const size_t nl = r.nl, ld = r.ld, per = nl * ld, nchol = r.nchol;
  double* mt = static_cast<double*>(wsbuf(c, WS_RK_MT, 2 * 5 * per * sizeof(double)));
  hipLaunchKernelGGL(k_rkb_mt, dim3(grid_for(5 * per)), dim3(256), 0, c.stream, dp, nl, ld, mt);
  double* t = static_cast<double*>(wsbuf(c, WS_RK_T, std::max<size_t>(1, 2 * nchol) * sizeof(double)));
  double* J = static_cast<double*>(wsbuf(c, WS_RK_J, 2 * 5 * per * sizeof(double)));
  if (nchol) {
    hipLaunchKernelGGL(k_rkb_trace, dim3((unsigned)std::min<size_t>(nchol, 65535)), dim3(256), 0, c.stream, r.nat, nchol,
                       per, mt, t);
    const size_t bx = std::min<size_t>((5 * per + 255) / 256, 4096);
    size_t ns = std::max<size_t>(1, std::min<size_t>(nchol, 2048 / bx));
    const size_t lper = (nchol + ns - 1) / ns;
    ns = (nchol + lper - 1) / lper;
    double* jp = static_cast<double*>(wsbuf(c, WS_RK_JPART, ns * 2 * 5 * per * sizeof(double)));
    hipLaunchKernelGGL(k_rkb_j_part, dim3((unsigned)bx, (unsigned)ns), dim3(256), 0, c.stream, r.nat, nchol, per, t, lper,
                       jp);
    hipLaunchKernelGGL(k_rkb_j_reduce, dim3(grid_for(2 * 5 * per)), dim3(256), 0, c.stream, jp, (int)ns, 5 * per, J);
  } else {
    hipLaunchKernelGGL(k_zero3, dim3(grid_for(2 * 5 * per)), dim3(256), 0, c.stream, J, 2 * 5 * per);
  }
  RR_HIP(hipGetLastError());
  return J;
}

}  // namespace
}  // namespace rr

using rr::Context;
using rr::Error;
using rr::guarded;

extern "C" {

int rr_rkb_cholesky_upload(rr_handle h, int set, size_t n_large, size_t n_chol, const double* large,
                           const double* small0, const double* small1, const double* cross_fwd,
                           const double* cross_bwd) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(set >= 0 && set < rr::kRkbSets, RR_ERR_ARG, "RKB Cholesky set out of range");
    RR_REQUIRE(large && small0 && small1 && cross_fwd && cross_bwd, RR_ERR_ARG, "RkbCholesky: null vectors");
    rr::RkbChol& r = c.rkb[set];
    RR_HIP(hipStreamSynchronize(c.stream));
    if (r.nat) RR_HIP(hipFree(r.nat));
    if (r.tr) RR_HIP(hipFree(r.tr));
    r = rr::RkbChol();
    const size_t nl = n_large, ld = rr::roundup(std::max<size_t>(nl, 1), 8), per = nl * ld;
    RR_HIP(hipMalloc(&r.nat, std::max<size_t>(1, 5 * n_chol * per) * 2 * sizeof(double)));
    RR_HIP(hipMalloc(&r.tr, std::max<size_t>(1, 3 * n_chol * per) * 2 * sizeof(double)));
    r.nl = nl, r.nchol = n_chol, r.ld = ld;
    const double* src[5] = {large, small0, small1, cross_fwd, cross_bwd};
    const size_t lchunk = std::max<size_t>(1, std::min<size_t>(n_chol, (size_t(1) << 25) / std::max<size_t>(1, nl * nl)));
    for (int comp = 0; comp < 5; ++comp) {
      const int is_real = comp == rr::kLarge ? 1 : 0;
      const size_t cw = is_real ? 1 : 2;
      for (size_t l0 = 0; l0 < n_chol; l0 += lchunk) {
        const size_t lc = std::min(lchunk, n_chol - l0);
        const double* d = rr::in_buf(c, rr::WS_RK_STAGE, src[comp] + l0 * nl * nl * cw, lc * nl * nl * cw);
        double* nat = r.nat + 2 * ((size_t)comp * n_chol + l0) * per;
        double* tr = comp < 3 ? r.tr + 2 * ((size_t)comp * n_chol + l0) * per : nullptr;
        hipLaunchKernelGGL(rr::k_rkb_store, dim3(rr::grid_for(lc * per)), dim3(256), 0, c.stream, d, is_real, lc, nl, ld,
                           nat, tr);
        RR_HIP(hipGetLastError());
        RR_HIP(hipStreamSynchronize(c.stream));  // the staging buffer is reused by the next chunk
      }
    }
    // Hermiticity of the flavor components (large, small[0], small[1]), for the factor-form exchange
    const size_t cnt = 3 * n_chol * per;
    const int nb = rr::grid_for(cnt, 256, 1024);
    double* hd = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_HERM, 2 * (size_t)nb * sizeof(double)));
    hipLaunchKernelGGL(rr::k_rkb_herm, dim3(nb), dim3(256), 0, c.stream, r.nat, r.tr, cnt, hd);
    RR_HIP(hipGetLastError());
    std::vector<double> hh(2 * (size_t)nb);
    RR_HIP(hipMemcpyAsync(hh.data(), hd, hh.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
    double dmax = 0.0, mmax = 0.0;
    for (int b = 0; b < nb; ++b) dmax = std::max(dmax, hh[2 * b]), mmax = std::max(mmax, hh[2 * b + 1]);
    r.herm_dev = mmax > 0.0 ? dmax / mmax : 0.0;
    RR_CONTROL("rkb_herm_dev", r.herm_dev);
  });
}

int rr_rkb_fock_cholesky(rr_handle h, int set, size_t n, const double* h_rkb, const double* p, double* fock) {
  return guarded(h, [&](Context& c) {
    rr::RkbChol& r = rr::get_set(c, set);
    RR_REQUIRE(n % 4 == 0, RR_ERR_DIM, "rkbFockMatrix: RKB basis dimension must be a multiple of 4");
    RR_REQUIRE(n / 4 == r.nl, RR_ERR_DIM, "rkbFockMatrix: RKB Cholesky dimension does not match H_RKB");
    RR_REQUIRE(h_rkb && p && fock, RR_ERR_ARG, "rkbFockMatrix: null argument");
    const size_t nl = r.nl, ld = r.ld, per = nl * ld, nchol = r.nchol, N = n;
    const double* dp = rr::in_buf(c, rr::WS_RK_P, p, 2 * N * N);
    const double* dh = rr::in_buf(c, rr::WS_RK_H, h_rkb, 2 * N * N);
    double* J = rr::rkb_coulomb(c, r, dp);  // Coulomb: traces, then J = sum_L t_L B_L per component
    // Exchange K_fg = sum_L B^f_L P_fg B^g_L
    double* K = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_K, 2 * N * N * sizeof(double)));
    hipLaunchKernelGGL(rr::k_zero3, dim3(rr::grid_for(2 * N * N)), dim3(256), 0, c.stream, K, 2 * N * N);
    const size_t ld4 = 4 * ld;
    const size_t lchunk = std::max<size_t>(1, std::min<size_t>(nchol, (size_t(1) << 26) / std::max<size_t>(1, nl * ld4)));
    double* ps = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_PST, 2 * 4 * ld * ld * sizeof(double)));
    double* tt = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_TT, std::max<size_t>(1, 2 * lchunk * nl * ld4) * sizeof(double)));
    for (int g = 0; g < 4 && nchol; ++g) {
      hipLaunchKernelGGL(rr::k_rkb_pstack, dim3(rr::grid_for(4 * ld * ld)), dim3(256), 0, c.stream, dp, nl, ld, g, ps);
      RR_HIP(hipGetLastError());
      const int cg = rr::flavor_comp(g);
      for (size_t l0 = 0; l0 < nchol; l0 += lchunk) {
        const size_t lc = std::min(lchunk, nchol - l0);
        {  // Tt_L[b][(f,m)] = sum_s Bt^g_L[b][s] Pstack[(f,m)][s]
          rr::GemmArgs ga;
          ga.A = r.tr + 2 * ((size_t)cg * nchol + l0) * per, ga.lda = (long long)ld, ga.sAb = (long long)per;
          ga.B = ps, ga.ldb = (long long)ld, ga.sBb = 0;
          ga.C = tt, ga.ldcm = (long long)ld4, ga.ldcn = 1, ga.sCb = (long long)(nl * ld4);
          ga.M = (int)nl, ga.N = (int)ld4, ga.K = (int)ld, ga.nbatch = (int)lc;
          rr::GemmSpec sp;
          sp.cplx = true, sp.cfg = 1;
          rr::run_gemm(c, sp, ga);
        }
        for (int f = 0; f < 4; ++f) {  // K_fg^T[b][a] += sum_L sum_m Tt_L[b][(f,m)] B^f_L[a][m]
          const int cf = rr::flavor_comp(f);
          const size_t ktot = lc * ld;
          const int nsplit0 = rr::splits_for(nl, nl, ktot);
          const size_t kper = rr::roundup((ktot + nsplit0 - 1) / nsplit0, 16);
          const int nsplit = (int)((ktot + kper - 1) / kper);
          double* part = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_KPART, (size_t)nsplit * 2 * nl * nl * sizeof(double)));
          rr::GemmArgs gk;
          gk.A = tt + 2 * f * ld, gk.lda = (long long)ld4;
          gk.B = r.nat + 2 * ((size_t)cf * nchol + l0) * per, gk.ldb = (long long)ld;
          gk.C = part, gk.ldcm = (long long)nl, gk.ldcn = 1, gk.sCz = (long long)(nl * nl);
          gk.M = (int)nl, gk.N = (int)nl, gk.K = (int)ld;
          gk.kbatch = (int)lc, gk.sAk = (long long)(nl * ld4), gk.sBk = (long long)per;
          gk.kper = (int)kper, gk.nsplit = nsplit;
          rr::GemmSpec sp;
          sp.cplx = true, sp.cfg = 1;
          rr::run_gemm(c, sp, gk);
          hipLaunchKernelGGL(rr::k_rkb_kadd, dim3(rr::grid_for(nl * nl)), dim3(256), 0, c.stream, part, nsplit, nl, f, g, K);
          RR_HIP(hipGetLastError());
        }
      }
    }
    rr::OutBuf of = rr::out_buf(c, rr::WS_RK_OUT, fock, 2 * N * N);
    hipLaunchKernelGGL(rr::k_rkb_assemble, dim3(rr::grid_for(N * N)), dim3(256), 0, c.stream, dh, J, K, nl, ld, of.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, of);
  });
}

int rr_rkb_fock_cholesky_factor(rr_handle h, int set, size_t n, const double* h_rkb, const double* s_full,
                                const double* x_full, size_t nx, const double* p, double* fock) {
  return guarded(h, [&](Context& c) {
    rr::RkbChol& r = rr::get_set(c, set);
    RR_REQUIRE(n % 4 == 0, RR_ERR_DIM, "rkbFockMatrix: RKB basis dimension must be a multiple of 4");
    RR_REQUIRE(n / 4 == r.nl, RR_ERR_DIM, "rkbFockMatrix: RKB Cholesky dimension does not match H_RKB");
    RR_REQUIRE(h_rkb && s_full && x_full && p && fock, RR_ERR_ARG, "rkbFockMatrix: null argument");
    RR_REQUIRE(nx > 0 && nx <= n, RR_ERR_DIM, "rkbFockMatrix: x_full column count out of range");
    RR_REQUIRE(r.herm_dev <= 1e-10, RR_ERR_ARG,
               "rkbFockMatrix (Cholesky factor form): the large/small flavor vectors are not Hermitian");
    const size_t nl = r.nl, ld = r.ld, per = nl * ld, nchol = r.nchol, N = n;
    const double* dp = rr::in_buf(c, rr::WS_RK_P, p, 2 * N * N);
    const double* dh = rr::in_buf(c, rr::WS_RK_H, h_rkb, 2 * N * N);
    const double* dS = rr::in_buf(c, rr::WS_RK_S, s_full, 2 * N * N);
    const double* dX = rr::in_buf(c, rr::WS_RK_XF, x_full, 2 * N * nx);
    double* J = rr::rkb_coulomb(c, r, dp);
    // occupied factor F = C sqrt(occ) with F F^H = P: eigenvectors of X^H (S P S) X above 1/2
    const double* rows = nullptr;
    std::vector<double> occ;
    const size_t n_occ = rr::recover_occupied(c, true, N, nx, dS, dp, dX, 1.0, &rows, occ, true);
    double* K = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_K, 2 * N * N * sizeof(double)));
    if (n_occ && nchol) {
      double* ct = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_CT2, 2 * n_occ * N * sizeof(double)));
      rr::dgemm_small(c, true, 'N', 'T', n_occ, N, nx, rows, nx, dX, nx, ct, N);  // C^T
      const double* docc = rr::to_device<double>(c, rr::WS_RK_OCC, occ.data(), n_occ);
      const size_t rp = rr::roundup(n_occ, 8);
      double* fp = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_FP, 2 * 4 * rp * ld * sizeof(double)));
      hipLaunchKernelGGL(rr::k_rkb_fpack, dim3(rr::grid_for(4 * rp * ld)), dim3(256), 0, c.stream, ct, N, docc, n_occ, nl,
                         rp, ld, fp);
      RR_HIP(hipGetLastError());
      // W[(f, a)][(L, i)] = (B^f_L F_f)[a][i] for a chunk of vectors (at most 4 GiB of W), then the
      // lower tiles of W W^H into split partial sums; every chunk adds its own partial sums
      const size_t wl = N * rp * 2 * sizeof(double);
      const size_t lchunk = std::max<size_t>(1, std::min<size_t>(nchol, (size_t(4) << 30) / wl));
      std::vector<int> nsplit;
      std::vector<size_t> kper;
      for (size_t l0 = 0; l0 < nchol; l0 += lchunk) {
        const size_t ktot = std::min(lchunk, nchol - l0) * rp;
        const int ns0 = rr::splits_for(N, N, ktot);
        const size_t kp = rr::roundup((ktot + ns0 - 1) / ns0, 16);
        kper.push_back(kp), nsplit.push_back((int)((ktot + kp - 1) / kp));
      }
      int nparts = 0;
      for (int v : nsplit) nparts += v;
      double* W = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_WF, lchunk * wl));
      double* part = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_KPART, (size_t)nparts * 2 * N * N * sizeof(double)));
      int zoff = 0;
      for (size_t l0 = 0, ci = 0; l0 < nchol; l0 += lchunk, ++ci) {
        const size_t lc = std::min(lchunk, nchol - l0), ldw = lc * rp;
        for (int f = 0; f < 4; ++f) {
          rr::GemmArgs g;
          g.A = r.nat + 2 * ((size_t)rr::flavor_comp(f) * nchol + l0) * per, g.lda = (long long)ld, g.sAb = (long long)per;
          g.B = fp + 2 * (size_t)f * rp * ld, g.ldb = (long long)ld, g.sBb = 0;
          g.C = W + 2 * (size_t)f * nl * ldw, g.ldcm = (long long)ldw, g.ldcn = 1, g.sCb = (long long)rp;
          g.M = (int)nl, g.N = (int)rp, g.K = (int)ld, g.nbatch = (int)lc;
          rr::GemmSpec sp;
          sp.cplx = true, sp.cfg = 1;
          rr::run_gemm(c, sp, g);
        }
        rr::GemmArgs g;
        g.A = W, g.lda = (long long)ldw, g.B = W, g.ldb = (long long)ldw;
        g.C = part + 2 * (size_t)zoff * N * N, g.ldcm = (long long)N, g.ldcn = 1, g.sCz = (long long)(N * N);
        g.M = (int)N, g.N = (int)N, g.K = (int)ldw, g.kper = (int)kper[ci], g.nsplit = nsplit[ci];
        g.tri = 1, g.conj = 2;
        rr::GemmSpec sp;
        sp.cplx = true, sp.cfg = 1;
        rr::run_gemm(c, sp, g);
        zoff += nsplit[ci];
      }
      hipLaunchKernelGGL(rr::k_rkb_kherm, dim3(rr::grid_for(N * N)), dim3(256), 0, c.stream, part, nparts, N, K);
    } else {
      hipLaunchKernelGGL(rr::k_zero3, dim3(rr::grid_for(2 * N * N)), dim3(256), 0, c.stream, K, 2 * N * N);
    }
    RR_HIP(hipGetLastError());
    rr::OutBuf of = rr::out_buf(c, rr::WS_RK_OUT, fock, 2 * N * N);
    hipLaunchKernelGGL(rr::k_rkb_assemble, dim3(rr::grid_for(N * N)), dim3(256), 0, c.stream, dh, J, K, nl, ld, of.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, of);
  });
}

int rr_rkb_cholesky_to_mo(rr_handle h, int set, const double* c_mo, size_t n_mo, size_t n_negative, double threshold,
                          double* out, size_t capacity, size_t* n_out) {
  return guarded(h, [&](Context& c) {
    rr::RkbChol& r = rr::get_set(c, set);
    RR_REQUIRE(c_mo && out && n_out, RR_ERR_ARG, "rkbCholeskyToMo: null argument");
    RR_REQUIRE(n_negative <= n_mo, RR_ERR_ARG, "rkbCholeskyToMo: n_negative exceeds the MO count");
    *n_out = 0;
    const size_t nl = r.nl, ld = r.ld, per = nl * ld, nchol = r.nchol;
    const size_t nn = n_negative, m = n_mo - nn, ldm2 = rr::roundup(std::max<size_t>(m * m, 1), 8);
    const double* dc = rr::in_buf(c, rr::WS_RK_C, c_mo, 2 * 4 * nl * n_mo);
    // operands per flavor: ct (C^T rows) and ch (C^H rows), m x ld each
    double* ct = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_CT, std::max<size_t>(1, 2 * 4 * m * ld) * sizeof(double)));
    double* ch = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_CH, std::max<size_t>(1, 2 * 4 * m * ld) * sizeof(double)));
    for (int f = 0; f < 4; ++f)
      hipLaunchKernelGGL(rr::k_rkb_cpack, dim3(rr::grid_for(m * ld)), dim3(256), 0, c.stream, dc, nl, n_mo, nn, m, ld, f,
                         ct + 2 * f * m * ld, ch + 2 * f * m * ld);
    RR_HIP(hipGetLastError());
    // A[L][(q,p)] = W_L entries (transposed bp), rows of length ldm2
    double* A = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_A, std::max<size_t>(1, 2 * nchol * ldm2) * sizeof(double)));
    hipLaunchKernelGGL(rr::k_zero3, dim3(rr::grid_for(2 * nchol * ldm2)), dim3(256), 0, c.stream, A, 2 * nchol * ldm2);
    const size_t lchunk = std::max<size_t>(1, std::min<size_t>(nchol, (size_t(1) << 26) / std::max<size_t>(1, m * ld)));
    double* X = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_X, std::max<size_t>(1, 2 * lchunk * m * ld) * sizeof(double)));
    hipLaunchKernelGGL(rr::k_zero3, dim3(rr::grid_for(2 * lchunk * m * ld)), dim3(256), 0, c.stream, X, 2 * lchunk * m * ld);
    // (bra flavor, ket flavor, component): C_bra^H B C_ket
    const int terms[6][3] = {{0, 0, rr::kLarge}, {1, 1, rr::kLarge}, {2, 2, rr::kSmall0},
                             {3, 3, rr::kSmall1}, {2, 3, rr::kCrossF}, {3, 2, rr::kCrossB}};
    for (size_t l0 = 0; l0 < nchol && m > 0; l0 += lchunk) {
      const size_t lc = std::min(lchunk, nchol - l0);
      for (int ti = 0; ti < 6; ++ti) {
        const int fb = terms[ti][0], fk = terms[ti][1], comp = terms[ti][2];
        {  // X_L[q][a] = sum_mm C_ket^T[q][mm] B_L[a][mm]
          rr::GemmArgs g;
          g.A = ct + 2 * fk * m * ld, g.lda = (long long)ld, g.sAb = 0;
          g.B = r.nat + 2 * ((size_t)comp * nchol + l0) * per, g.ldb = (long long)ld, g.sBb = (long long)per;
          g.C = X, g.ldcm = (long long)ld, g.ldcn = 1, g.sCb = (long long)(m * ld);
          g.M = (int)m, g.N = (int)nl, g.K = (int)ld, g.nbatch = (int)lc;
          rr::GemmSpec sp;
          sp.cplx = true, sp.cfg = 1;
          rr::run_gemm(c, sp, g);
        }
        {  // W_L^T[q][p] (+)= sum_a C_bra^H[p][a] X_L[q][a]
          rr::GemmArgs g;
          g.A = ch + 2 * fb * m * ld, g.lda = (long long)ld, g.sAb = 0;
          g.B = X, g.ldb = (long long)ld, g.sBb = (long long)(m * ld);
          g.C = A + 2 * l0 * ldm2, g.ldcm = 1, g.ldcn = (long long)m, g.sCb = (long long)ldm2;
          g.M = (int)m, g.N = (int)m, g.K = (int)ld, g.nbatch = (int)lc;
          g.beta = ti == 0 ? 0.0 : 1.0;
          rr::GemmSpec sp;
          sp.cplx = true, sp.cfg = 1;
          rr::run_gemm(c, sp, g);
        }
      }
    }
    if (nn == 0) {
      RR_REQUIRE(capacity >= nchol, RR_ERR_ARG, "rkbCholeskyToMo: output capacity too small");
      if (nchol && m)
        RR_HIP(hipMemcpy2DAsync(out, 2 * m * m * sizeof(double), A, 2 * ldm2 * sizeof(double), 2 * m * m * sizeof(double),
                                nchol, hipMemcpyDefault, c.stream));
      *n_out = nchol;
      return;
    }
    // Recompression (RkbCholesky.cpp:244-283): G = A A^H, keep eigenvalues > threshold / nchol,
    // A'_k = sum_L conj(u_k(L)) A_L, embedded in nmo x nmo.
    RR_REQUIRE(nchol > 0, RR_ERR_SOLVER, "recompressMoVectors: no Cholesky direction survives the compression");
    double* G = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_G, 2 * nchol * nchol * sizeof(double)));
    {
      rr::GemmArgs g;
      g.A = A, g.lda = (long long)ldm2;
      g.B = A, g.ldb = (long long)ldm2;
      g.C = G, g.ldcm = (long long)nchol, g.ldcn = 1;
      g.M = (int)nchol, g.N = (int)nchol, g.K = (int)ldm2, g.conj = 2;
      rr::GemmSpec sp;
      sp.cplx = true, sp.cfg = 1;
      rr::run_gemm(c, sp, g);
    }
    double* w = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_W, nchol * sizeof(double)));
    rr::eig_device(c, true, nchol, G, w);  // row j = eigenvector j, ascending
    std::vector<double> wh(nchol);
    RR_HIP(hipMemcpyAsync(wh.data(), w, nchol * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
    const double tau = threshold / static_cast<double>(std::max<size_t>(1, nchol));
    size_t first = 0;
    while (first < nchol && !(wh[first] > tau)) ++first;
    const size_t nk = nchol - first;
    RR_REQUIRE(nk > 0, RR_ERR_SOLVER, "recompressMoVectors: no Cholesky direction survives the compression");
    RR_REQUIRE(capacity >= nk, RR_ERR_ARG, "rkbCholeskyToMo: output capacity too small");
    const size_t ldl = rr::roundup(nchol, 8);
    double* U = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_U, 2 * nk * ldl * sizeof(double)));
    hipLaunchKernelGGL(rr::k_rkb_urows, dim3(rr::grid_for(nk * ldl)), dim3(256), 0, c.stream, G, nchol, first, nk, ldl, U);
    double* At = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_AT, 2 * m * m * ldl * sizeof(double)));
    hipLaunchKernelGGL(rr::k_rkb_transpose, dim3(rr::grid_for(m * m * ldl)), dim3(256), 0, c.stream, A, nchol, ldm2, m * m,
                       ldl, At);
    double* Ap = static_cast<double*>(rr::wsbuf(c, rr::WS_RK_AP, 2 * nk * m * m * sizeof(double)));
    {  // A'[k][j] = sum_L conj(U[k][L]) At[j][L]
      rr::GemmArgs g;
      g.A = U, g.lda = (long long)ldl;
      g.B = At, g.ldb = (long long)ldl;
      g.C = Ap, g.ldcm = (long long)(m * m), g.ldcn = 1;
      g.M = (int)nk, g.N = (int)(m * m), g.K = (int)ldl, g.conj = 1;
      rr::GemmSpec sp;
      sp.cplx = true, sp.cfg = 1;
      rr::run_gemm(c, sp, g);
    }
    rr::OutBuf ob = rr::out_buf(c, rr::WS_RK_OUT, out, 2 * nk * n_mo * n_mo);
    hipLaunchKernelGGL(rr::k_rkb_embed, dim3(rr::grid_for(nk * n_mo * n_mo)), dim3(256), 0, c.stream, Ap, nk, m, m * m, nn,
                       n_mo, ob.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, ob);
    *n_out = nk;
  });
}

int rr_rkb_cholesky_release(rr_handle h, int set) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(set >= 0 && set < rr::kRkbSets, RR_ERR_ARG, "RKB Cholesky set out of range");
    rr::RkbChol& r = c.rkb[set];
    RR_HIP(hipStreamSynchronize(c.stream));
    if (r.nat) RR_HIP(hipFree(r.nat));
    if (r.tr) RR_HIP(hipFree(r.tr));
    r = rr::RkbChol();
  });
}

}  // extern "C"
