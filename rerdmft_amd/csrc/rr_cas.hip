// librerdmft_gpu, the CASSCF stage's RI contractions on the device (src/MCSCF/CasFock.cpp:70-150,
// RI backends :244-262). The MO tensor B (n_aux x n^2, row P = the n x n matrix B_P, real or
// interleaved complex, Hermitian B_P as every RI MO tensor of the reference is) defines
// <ab|cd> = sum_P B(P, a n + c) B(P, b n + d) (no conjugation).
//
//   two-electron operator  G[rho]_rp = sum_st rho_st (<sr|tp> - <rs|tp>) = J - K,
//       J = sum_P c_P B_P with c_P = sum_st rho_st B_P(s,t);
//       K = sum_P B_P rho^T B_P evaluated through rho = U diag(w) U^H (|w| > 1e-14 kept, as the
//       reference's lowRankFactors): Y_P = B_P conj(U), K = sum_(P,k) Y_P(r,k) w_k conj(Y_P(p,k)).
//   active integrals       (tu|vw) = sum_P B(P, t n + u) B(P, v n + w) over the active indices.
//   active cumulant Fock   F(p,q) = sum_r sum_P B(P, act_r n + p) T_P(r,q),
//                          T_P(r,q) = sum_(s,t) B(P, act_s n + act_t) lambda[s,t,r,q].
#include <algorithm>
#include <cmath>

#include "rr_core.hpp"

namespace {

using rr::Context;
using rr::Error;

enum CasWs {
  WS_CS_B = rr::WS_SIGMA_BASE + 6000, WS_CS_RHO, WS_CS_C, WS_CS_JV, WS_CS_EIG, WS_CS_W, WS_CS_UH, WS_CS_Y, WS_CS_YW,
  WS_CS_KP, WS_CS_K, WS_CS_IDX, WS_CS_BA, WS_CS_LAM, WS_CS_TM, WS_CS_X, WS_CS_YT, WS_CS_OUT, WS_CS_WK
};

// c[P] = sum_k B[P][k] rho[k] (k over n^2; complex: plain products). One workgroup per P.
__global__ __launch_bounds__(256) void k_cas_c(const double* b, const double* rho, size_t n2, int cw, double* c) {
  __shared__ double sr[256], si[256];
  const double* row = b + blockIdx.x * n2 * cw;
  double ar = 0.0, ai = 0.0;
  for (size_t k = threadIdx.x; k < n2; k += blockDim.x) {
    if (cw == 1) {
      ar += row[k] * rho[k];
    } else {
      const double br = row[2 * k], bi = row[2 * k + 1], pr = rho[2 * k], pi = rho[2 * k + 1];
      ar += br * pr - bi * pi;
      ai += br * pi + bi * pr;
    }
  }
  sr[threadIdx.x] = ar, si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sr[threadIdx.x] += sr[threadIdx.x + s], si[threadIdx.x] += si[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    c[blockIdx.x * cw] = sr[0];
    if (cw == 2) c[blockIdx.x * 2 + 1] = si[0];
  }
}

// jv[k] = sum_P c[P] B[P][k]
__global__ void k_cas_j(const double* b, const double* c, size_t naux, size_t n2, int cw, double* jv) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n2; k += (size_t)gridDim.x * blockDim.x) {
    double ar = 0.0, ai = 0.0;
    for (size_t P = 0; P < naux; ++P) {
      if (cw == 1) {
        ar += c[P] * b[P * n2 + k];
      } else {
        const double cr = c[2 * P], ci = c[2 * P + 1], br = b[2 * (P * n2 + k)], bi = b[2 * (P * n2 + k) + 1];
        ar += cr * br - ci * bi;
        ai += cr * bi + ci * br;
      }
    }
    jv[k * cw] = ar;
    if (cw == 2) jv[k * 2 + 1] = ai;
  }
}

// uh[j][t] = conj(evec[idx[j]][t]) for t < n, zero for t in [n, ld) (rows of the eigensolver output
// are eigenvectors); wk[j] = w[idx[j]].
__global__ void k_cas_uh(const double* evec, const double* w, const int* idx, size_t nk, size_t n, size_t ld, int cw,
                         double* uh, double* wk) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nk * ld; i += (size_t)gridDim.x * blockDim.x) {
    const size_t j = i / ld, t = i % ld;
    const size_t src = (size_t)idx[j] * n + t;
    for (int q = 0; q < cw; ++q) uh[i * cw + q] = t < n ? (q ? -evec[src * cw + q] : evec[src * cw + q]) : 0.0;
    if (t == 0) wk[j] = w[idx[j]];
  }
}

// yw[row][k] = y[row][k] * wk[k] (ld columns, the padding stays zero)
__global__ void k_cas_scale_cols(const double* y, const double* wk, size_t rows, size_t nk, size_t ld, int cw,
                                 double* yw) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < rows * ld; i += (size_t)gridDim.x * blockDim.x) {
    const size_t k = i % ld;
    const double s = k < nk ? wk[k] : 0.0;
    for (int q = 0; q < cw; ++q) yw[i * cw + q] = y[i * cw + q] * s;
  }
}

__global__ void k_cas_zero(double* p, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// g = jv - sum_s kpart[s]
__global__ void k_cas_g(const double* jv, const double* kpart, int nsplit, size_t count, double* g) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    double k = 0.0;
    for (int s = 0; s < nsplit; ++s) k += kpart[(size_t)s * count + i];
    g[i] = jv[i] - k;
  }
}

// out[P][(s,t)] = B[P][act[s] n + act[t]]
__global__ void k_cas_gather_act(const double* b, const long long* act, size_t naux, size_t n, size_t a, int cw,
                                 double* out) {
  const size_t a2 = a * a;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < naux * a2; i += (size_t)gridDim.x * blockDim.x) {
    const size_t P = i / a2, st = i % a2, s = st / a, t = st % a;
    const size_t src = P * n * n + (size_t)act[s] * n + (size_t)act[t];
    for (int q = 0; q < cw; ++q) out[i * cw + q] = b[src * cw + q];
  }
}

// x[(r,P)][p] = B[P][act[r] n + p];  yt[(r,P)][q] = tm[P][r a + q]
__global__ void k_cas_gather_fock(const double* b, const double* tm, const long long* act, size_t naux, size_t n,
                                  size_t a, int cw, double* x, double* yt) {
  const size_t rows = a * naux;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < rows * n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t row = i / n, p = i % n, r = row / naux, P = row % naux;
    const size_t src = P * n * n + (size_t)act[r] * n + p;
    for (int q = 0; q < cw; ++q) x[i * cw + q] = b[src * cw + q];
    if (p < a)
      for (int q = 0; q < cw; ++q) yt[(row * a + p) * cw + q] = tm[(P * a * a + r * a + p) * cw + q];
  }
}

// out[P][(i,j)] = B[P][I[i] n + J[j]]
__global__ void k_cc_gather_pairs(const double* b, const long long* I, size_t ni, const long long* J, size_t nj,
                                  size_t naux, size_t n, int cw, double* out) {
  const size_t m = ni * nj;
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < naux * m; x += (size_t)gridDim.x * blockDim.x) {
    const size_t P = x / m, ij = x % m, i = ij / nj, j = ij % nj;
    const size_t src = P * n * n + (size_t)I[i] * n + (size_t)J[j];
    for (int q = 0; q < cw; ++q) out[x * cw + q] = b[src * cw + q];
  }
}

// out[p][q][r][s] = x[(p r)][(q s)]
__global__ void k_cc_permute(const double* x, size_t np, size_t nq, size_t nr, size_t ns, int cw, double* out) {
  const size_t total = np * nq * nr * ns;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < total; o += (size_t)gridDim.x * blockDim.x) {
    const size_t s_ = o % ns, r = (o / ns) % nr, q = (o / (ns * nr)) % nq, p = o / (ns * nr * nq);
    const size_t src = (p * nr + r) * (nq * ns) + q * ns + s_;
    for (int k = 0; k < cw; ++k) out[o * cw + k] = x[src * cw + k];
  }
}

const long long* upload_act(Context& c, size_t a, const size_t* act, size_t n, int slot = 0,
                            const char* what = "CAS: active index out of range") {
  std::vector<long long> h(a);
  for (size_t i = 0; i < a; ++i) {
    RR_REQUIRE(act[i] < n, RR_ERR_ARG, what);
    h[i] = (long long)act[i];
  }
  long long* d = static_cast<long long*>(rr::wsbuf(c, WS_CS_IDX + 1000 * slot, std::max<size_t>(a, 1) * sizeof(long long)));
  RR_HIP(hipMemcpyAsync(d, h.data(), a * sizeof(long long), hipMemcpyHostToDevice, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));  // h is released on return
  return d;
}

}  // namespace

using rr::guarded;

extern "C" {

int rr_cas_two_electron_operator_ri(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b,
                                    const double* rho, double* g) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b && rho && g, RR_ERR_ARG, "casTwoElectronOperator: null input");
    if (n == 0) return;
    const int cw = is_complex ? 2 : 1;
    const bool cx = is_complex != 0;
    const size_t n2 = n * n;
    const double* dB = rr::in_buf(c, WS_CS_B, b, n_aux * n2 * cw);
    const double* dR = rr::in_buf(c, WS_CS_RHO, rho, n2 * cw);
    // Coulomb
    double* cvec = static_cast<double*>(rr::wsbuf(c, WS_CS_C, std::max<size_t>(n_aux, 1) * cw * sizeof(double)));
    double* jv = static_cast<double*>(rr::wsbuf(c, WS_CS_JV, n2 * cw * sizeof(double)));
    if (n_aux) hipLaunchKernelGGL(k_cas_c, dim3((unsigned)n_aux), dim3(256), 0, c.stream, dB, dR, n2, cw, cvec);
    hipLaunchKernelGGL(k_cas_j, dim3(rr::grid_for(n2)), dim3(256), 0, c.stream, dB, cvec, n_aux, n2, cw, jv);
    RR_HIP(hipGetLastError());
    // rho = U diag(w) U^H, |w| > 1e-14 kept (CasFock.cpp:45-68)
    double* ev = static_cast<double*>(rr::wsbuf(c, WS_CS_EIG, n2 * cw * sizeof(double)));
    double* w = static_cast<double*>(rr::wsbuf(c, WS_CS_W, n * sizeof(double)));
    RR_HIP(hipMemcpyAsync(ev, dR, n2 * cw * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    rr::eig_device(c, cx, n, ev, w);
    std::vector<double> wh(n);
    RR_HIP(hipMemcpyAsync(wh.data(), w, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
    std::vector<int> keep;
    for (size_t k = 0; k < n; ++k)
      if (std::fabs(wh[k]) > 1e-14) keep.push_back((int)k);
    const size_t nk = keep.size();
    rr::OutBuf og = rr::out_buf(c, WS_CS_OUT, g, n2 * cw);
    if (nk == 0 || n_aux == 0) {
      RR_HIP(hipMemcpyAsync(og.dev, jv, n2 * cw * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
      rr::finish_out(c, og);
      return;
    }
    const size_t ld = rr::roundup(n, 8), nkp = rr::roundup(nk, 8);
    int* didx = static_cast<int*>(rr::wsbuf(c, WS_CS_IDX + 100, nk * sizeof(int)));
    RR_HIP(hipMemcpyAsync(didx, keep.data(), nk * sizeof(int), hipMemcpyHostToDevice, c.stream));
    double* uh = static_cast<double*>(rr::wsbuf(c, WS_CS_UH, nk * ld * cw * sizeof(double)));
    double* wk = static_cast<double*>(rr::wsbuf(c, WS_CS_WK, nk * sizeof(double)));
    hipLaunchKernelGGL(k_cas_uh, dim3(rr::grid_for(nk * ld)), dim3(256), 0, c.stream, ev, w, didx, nk, n, ld, cw, uh, wk);
    RR_HIP(hipGetLastError());
    // Y[(P,r)][k] = sum_t B_P(r,t) conj(U)(t,k): one product over the n_aux n rows of B
    double* y = static_cast<double*>(rr::wsbuf(c, WS_CS_Y, n_aux * n * nkp * cw * sizeof(double)));
    double* yw = static_cast<double*>(rr::wsbuf(c, WS_CS_YW, n_aux * n * nkp * cw * sizeof(double)));
    hipLaunchKernelGGL(k_cas_zero, dim3(rr::grid_for(n_aux * n * nkp * cw)), dim3(256), 0, c.stream, y,
                       n_aux * n * nkp * cw);
    RR_HIP(hipGetLastError());
    rr::dgemm_small(c, cx, 'N', 'T', n_aux * n, nk, n, dB, n, uh, ld, y, nkp);
    hipLaunchKernelGGL(k_cas_scale_cols, dim3(rr::grid_for(n_aux * n * nkp)), dim3(256), 0, c.stream, y, wk, n_aux * n,
                       nk, nkp, cw, yw);
    RR_HIP(hipGetLastError());
    // K(r,p) = sum_(P,k) (w_k Y[(P,r)][k]) conj(Y[(P,p)][k]), a k-batched product over P, split for parallelism
    const size_t kflat = n_aux * nkp;
    // about 2048 workgroups over the output tiles, at least 128 k per split
    const size_t tiles = ((n + 63) / 64) * ((n + 63) / 64);
    const size_t ns = std::max<size_t>(1, std::min<size_t>(std::min<size_t>(256, (2048 + tiles - 1) / tiles),
                                                            std::max<size_t>(1, kflat / 128)));
    const size_t kper = rr::roundup((kflat + ns - 1) / ns, nkp);  // whole aux rows per split
    const int nsplit = (int)((kflat + kper - 1) / kper);
    double* kpart = static_cast<double*>(rr::wsbuf(c, WS_CS_KP, (size_t)nsplit * n2 * cw * sizeof(double)));
    rr::GemmArgs ga;
    ga.A = yw, ga.lda = nkp, ga.sAk = n * nkp;
    ga.B = y, ga.ldb = nkp, ga.sBk = n * nkp;
    ga.C = kpart, ga.ldcm = n, ga.ldcn = 1, ga.sCz = n2;
    ga.M = (int)n, ga.N = (int)n, ga.K = (int)nkp, ga.kbatch = (int)n_aux, ga.kper = (int)kper, ga.nsplit = nsplit;
    ga.conj = cx ? 2 : 0;
    rr::GemmSpec sp;
    sp.prec = rr::P64, sp.out = rr::P64, sp.cplx = cx, sp.cfg = n >= 256 ? 2 : 1;
    rr::run_gemm(c, sp, ga);
    hipLaunchKernelGGL(k_cas_g, dim3(rr::grid_for(n2 * cw)), dim3(256), 0, c.stream, jv, kpart, nsplit, n2 * cw, og.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, og);
  });
}

int rr_cas_active_integrals_ri(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b, size_t n_act,
                               const size_t* act, double* out) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b && out && (act || n_act == 0), RR_ERR_ARG, "casActiveIntegrals: null input");
    const int cw = is_complex ? 2 : 1;
    const size_t a2 = n_act * n_act;
    if (a2 == 0) return;
    const long long* dact = upload_act(c, n_act, act, n);
    const double* dB = rr::in_buf(c, WS_CS_B, b, n_aux * n * n * cw);
    double* ba = static_cast<double*>(rr::wsbuf(c, WS_CS_BA, std::max<size_t>(n_aux, 1) * a2 * cw * sizeof(double)));
    hipLaunchKernelGGL(k_cas_gather_act, dim3(rr::grid_for(n_aux * a2)), dim3(256), 0, c.stream, dB, dact, n_aux, n, n_act,
                       cw, ba);
    RR_HIP(hipGetLastError());
    rr::OutBuf oo = rr::out_buf(c, WS_CS_OUT, out, a2 * a2 * cw);
    if (n_aux == 0)
      hipLaunchKernelGGL(k_cas_zero, dim3(rr::grid_for(a2 * a2 * cw)), dim3(256), 0, c.stream, oo.dev, a2 * a2 * cw);
    else
      rr::dgemm_small(c, is_complex != 0, 'T', 'N', a2, a2, n_aux, ba, a2, ba, a2, oo.dev, a2);
    rr::finish_out(c, oo);
  });
}

int rr_cas_active_cumulant_fock_ri(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b, size_t n_act,
                                   const size_t* act, const double* lambda, double* f) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b && f && (n_act == 0 || (act && lambda)), RR_ERR_ARG, "casActiveCumulantFock: null input");
    const int cw = is_complex ? 2 : 1;
    const bool cx = is_complex != 0;
    const size_t a = n_act, a2 = a * a;
    rr::OutBuf of = rr::out_buf(c, WS_CS_OUT, f, std::max<size_t>(n * a * cw, 1));
    if (a == 0 || n == 0) {
      rr::finish_out(c, of);
      return;
    }
    if (n_aux == 0) {
      hipLaunchKernelGGL(k_cas_zero, dim3(rr::grid_for(n * a * cw)), dim3(256), 0, c.stream, of.dev, n * a * cw);
      rr::finish_out(c, of);
      return;
    }
    RR_REQUIRE(n >= a, RR_ERR_DIM, "casActiveCumulantFock: more active orbitals than orbitals");
    const long long* dact = upload_act(c, a, act, n);
    const double* dB = rr::in_buf(c, WS_CS_B, b, n_aux * n * n * cw);
    const double* dL = rr::in_buf(c, WS_CS_LAM, lambda, a2 * a2 * cw);
    double* bact = static_cast<double*>(rr::wsbuf(c, WS_CS_BA, n_aux * a2 * cw * sizeof(double)));
    hipLaunchKernelGGL(k_cas_gather_act, dim3(rr::grid_for(n_aux * a2)), dim3(256), 0, c.stream, dB, dact, n_aux, n, a, cw,
                       bact);
    RR_HIP(hipGetLastError());
    double* tm = static_cast<double*>(rr::wsbuf(c, WS_CS_TM, n_aux * a2 * cw * sizeof(double)));
    rr::dgemm_small(c, cx, 'N', 'N', n_aux, a2, a2, bact, a2, dL, a2, tm, a2);  // T[P][(r,q)]
    double* x = static_cast<double*>(rr::wsbuf(c, WS_CS_X, a * n_aux * n * cw * sizeof(double)));
    double* yt = static_cast<double*>(rr::wsbuf(c, WS_CS_YT, a * n_aux * a * cw * sizeof(double)));
    hipLaunchKernelGGL(k_cas_gather_fock, dim3(rr::grid_for(a * n_aux * n)), dim3(256), 0, c.stream, dB, tm, dact, n_aux, n,
                       a, cw, x, yt);
    RR_HIP(hipGetLastError());
    rr::dgemm_small(c, cx, 'T', 'N', n, a, a * n_aux, x, n, yt, a, of.dev, a);  // F = X^T Yt
    rr::finish_out(c, of);
  });
}

int rr_cc_physics_block(rr_handle h, int is_complex, size_t n_aux, size_t n, const double* b, size_t np, const size_t* p,
                        size_t nq, const size_t* q, size_t nr, const size_t* r, size_t ns, const size_t* s, double* out) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b && out, RR_ERR_ARG, "FactoredCcIntegrals::physicsBlock: null input");
    const int cw = is_complex ? 2 : 1;
    const size_t M = np * nr, N = nq * ns;
    if (M == 0 || N == 0) return;
    const char* msg = "FactoredCcIntegrals::physicsBlock: index out of range";
    const long long* dp = upload_act(c, np, p, n, 1, msg);
    const long long* dq = upload_act(c, nq, q, n, 2, msg);
    const long long* drr = upload_act(c, nr, r, n, 3, msg);
    const long long* ds = upload_act(c, ns, s, n, 4, msg);
    const double* dB = rr::in_buf(c, WS_CS_B, b, n_aux * n * n * cw);
    rr::OutBuf oo = rr::out_buf(c, WS_CS_OUT, out, M * N * cw);
    double* x = static_cast<double*>(rr::wsbuf(c, WS_CS_X, M * N * cw * sizeof(double)));
    if (n_aux == 0) {
      hipLaunchKernelGGL(k_cas_zero, dim3(rr::grid_for(M * N * cw)), dim3(256), 0, c.stream, x, M * N * cw);
    } else {
      double* ga = static_cast<double*>(rr::wsbuf(c, WS_CS_BA, n_aux * M * cw * sizeof(double)));
      double* gb = static_cast<double*>(rr::wsbuf(c, WS_CS_YT, n_aux * N * cw * sizeof(double)));
      hipLaunchKernelGGL(k_cc_gather_pairs, dim3(rr::grid_for(n_aux * M)), dim3(256), 0, c.stream, dB, dp, np, drr, nr,
                         n_aux, n, cw, ga);
      hipLaunchKernelGGL(k_cc_gather_pairs, dim3(rr::grid_for(n_aux * N)), dim3(256), 0, c.stream, dB, dq, nq, ds, ns,
                         n_aux, n, cw, gb);
      RR_HIP(hipGetLastError());
      rr::dgemm_small(c, is_complex != 0, 'T', 'N', M, N, n_aux, ga, M, gb, N, x, N);  // x = A^T B (no conjugation)
    }
    hipLaunchKernelGGL(k_cc_permute, dim3(rr::grid_for(M * N)), dim3(256), 0, c.stream, x, np, nq, nr, ns, cw, oo.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, oo);
  });
}

}  // extern "C"
