// librerdmft_gpu, one-time RI tensor preparation: the auxiliary-metric contraction of the raw
// three-centre integrals, the canonical orthogonalization of the auxiliary metric, and the AO-leg
// transform of an RI tensor (Cartesian -> spherical / any rectangular T). FP64 throughout (the
// integral side of the method stays FP64).
//
// All products are NT GEMMs of rr_gemm.hpp on padded operands:
//   metric contraction  B[k][j] = sum_r X[r][k] raw[r][j]       A = X^T (packed), B = raw^T (tiled transpose, chunked over j)
//   AO legs             W_P = B_P T  (stored transposed),  out_P = T^T W_P   (two batched GEMMs per chunk of P)
#include <rocsolver/rocsolver.h>

#include "rr_core.hpp"

namespace rr {
namespace {

enum TrWs {
  WS_TR_RAWT = WS_SIGMA_BASE + 2000, WS_TR_XT, WS_TR_IN, WS_TR_BPAD, WS_TR_TT, WS_TR_WT, WS_TR_OUT, WS_TR_S,
  WS_TR_W, WS_TR_X, WS_TR_RAW
};

#define RR_GRID_LOOP(i, total) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (total); i += (size_t)gridDim.x * blockDim.x)

// out[j][r] (row length ldr, zero-padded) = in[r][j0 + j] for j < jc, r < rows; in has row length ld_in.
// 32 x 32 tiles through LDS so both the reads and the writes are coalesced.
__global__ __launch_bounds__(256) void k_transpose_cols(const double* in, size_t rows, size_t ld_in, size_t j0, size_t jc,
                                                        size_t ldr, double* out) {
  __shared__ double tile[32][33];
  const size_t tr = static_cast<size_t>(blockIdx.y) * 32;  // r tile
  const size_t tj = static_cast<size_t>(blockIdx.x) * 32;  // j tile
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = 0; k < 32; k += 8) {
    const size_t r = tr + ty + k, j = tj + tx;
    tile[ty + k][tx] = (r < rows && j < jc) ? in[r * ld_in + j0 + j] : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < 32; k += 8) {
    const size_t j = tj + ty + k, r = tr + tx;
    if (j < jc && r < ldr) out[j * ldr + r] = tile[tx][ty + k];
  }
}

// out[r][c] (row length ldo, zero-padded to rows_out x ldo) = trans ? in[c][r] : in[r][c], in row length ldi.
__global__ void k_pack2(const double* in, size_t ldi, int trans, size_t rows, size_t cols, size_t ldo, size_t rows_out,
                        double* out) {
  RR_GRID_LOOP(i, rows_out * ldo) {
    const size_t r = i / ldo, c = i % ldo;
    out[i] = (r < rows && c < cols) ? (trans ? in[c * ldi + r] : in[r * ldi + c]) : 0.0;
  }
}

__global__ void k_zero2(double* p, size_t count) {
  RR_GRID_LOOP(i, count) p[i] = 0.0;
}

// X(:, j) = U(:, keep_j) / sqrt(w_keep_j) from row-major eigenvectors (row j = eigenvector j):
// x[i][jj] = v[first + jj][i] / sqrt(w[first + jj]).
__global__ void k_canon_x(const double* v, const double* w, size_t n, size_t first, size_t neff, double* x) {
  RR_GRID_LOOP(idx, n * neff) {
    const size_t i = idx / neff, jj = idx % neff;
    x[idx] = v[(first + jj) * n + i] / sqrt(w[first + jj]);
  }
}
// Symmetric S^-1/2 = sum_k v_k v_k^T / sqrt(w_k): x[i][j] = sum_k v[k][i] v[k][j] / sqrt(w[k]) via the
// NT engine on Y[k][i] = v[k][i] / w_k^(1/4): x = Y^T Y (the caller runs the GEMM).
__global__ void k_quarter(const double* v, const double* w, size_t n, size_t ld, double* y) {
  RR_GRID_LOOP(idx, n * ld) {
    const size_t i = idx / ld, k = idx % ld;  // y_t[i][k] = v[k][i] / w_k^(1/4)
    y[idx] = k < n ? v[k * n + i] / sqrt(sqrt(w[k])) : 0.0;
  }
}

}  // namespace
}  // namespace rr

using rr::Context;
using rr::Error;
using rr::guarded;

extern "C" {

int rr_ri_three_center_tensor(rr_handle h, size_t n, size_t n_aux_raw, const double* raw, const double* metric_x,
                              size_t n_aux_kept, double* out) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(raw && metric_x && out, RR_ERR_ARG, "riThreeCenterTensor: null argument");
    const size_t n2 = n * n;
    const size_t ldr = rr::roundup(std::max<size_t>(n_aux_raw, 1), 8);
    const double* draw = rr::in_buf(c, rr::WS_TR_RAW, raw, n_aux_raw * n2);
    const double* dx = rr::in_buf(c, rr::WS_TR_X, metric_x, n_aux_raw * n_aux_kept);
    rr::OutBuf ob = rr::out_buf(c, rr::WS_TR_OUT, out, n_aux_kept * n2);
    if (n2 == 0 || n_aux_kept == 0) {
      rr::finish_out(c, ob);
      return;
    }
    // A = X^T, padded (n_aux_kept x ldr)
    double* xt = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_XT, n_aux_kept * ldr * sizeof(double)));
    hipLaunchKernelGGL(rr::k_pack2, dim3(rr::grid_for(n_aux_kept * ldr)), dim3(256), 0, c.stream, dx, n_aux_kept, 1,
                       n_aux_kept, n_aux_raw, ldr, n_aux_kept, xt);
    // raw^T in chunks of columns j
    const size_t jchunk = std::max<size_t>(64, std::min<size_t>(n2, (size_t(1) << 27) / ldr));
    double* rt = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_RAWT, std::min(jchunk, n2) * ldr * sizeof(double)));
    for (size_t j0 = 0; j0 < n2; j0 += jchunk) {
      const size_t jc = std::min(jchunk, n2 - j0);
      dim3 grid((unsigned)((jc + 31) / 32), (unsigned)((ldr + 31) / 32));
      hipLaunchKernelGGL(rr::k_transpose_cols, grid, dim3(256), 0, c.stream, draw, n_aux_raw, n2, j0, jc, ldr, rt);
      RR_HIP(hipGetLastError());
      rr::GemmArgs g;
      g.A = xt, g.lda = (long long)ldr;
      g.B = rt, g.ldb = (long long)ldr;
      g.C = ob.dev + j0, g.ldcm = (long long)n2, g.ldcn = 1;
      g.M = (int)n_aux_kept, g.N = (int)jc, g.K = (int)ldr;
      rr::GemmSpec sp;
      sp.cfg = 2;
      rr::gemm_auto(c, sp, g);
    }
    rr::finish_out(c, ob);
  });
}

int rr_transform_ri_three_center_ao_legs(rr_handle h, size_t n_aux, size_t n_ao, const double* eri3_in,
                                         const double* t, size_t n_out, double* eri3_out) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(eri3_in && t && eri3_out, RR_ERR_ARG, "transformRiThreeCenterAoLegs: null argument");
    const size_t ld = rr::roundup(std::max<size_t>(n_ao, 1), 8);
    const double* din = rr::in_buf(c, rr::WS_TR_IN, eri3_in, n_aux * n_ao * n_ao);
    const double* dt = rr::in_buf(c, rr::WS_TR_X, t, n_ao * n_out);
    rr::OutBuf ob = rr::out_buf(c, rr::WS_TR_OUT, eri3_out, n_aux * n_out * n_out);
    if (n_aux == 0 || n_out == 0 || n_ao == 0) {
      if (n_aux && n_out)
        hipLaunchKernelGGL(rr::k_zero2, dim3(rr::grid_for(n_aux * n_out * n_out)), dim3(256), 0, c.stream, ob.dev,
                           n_aux * n_out * n_out);
      rr::finish_out(c, ob);
      return;
    }
    // Tt[o][a] = T[a][o], padded
    double* tt = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_TT, n_out * ld * sizeof(double)));
    hipLaunchKernelGGL(rr::k_pack2, dim3(rr::grid_for(n_out * ld)), dim3(256), 0, c.stream, dt, n_out, 1, n_out, n_ao, ld,
                       n_out, tt);
    const size_t per = std::max(n_ao * ld, n_out * ld);
    const size_t pchunk = std::max<size_t>(1, std::min<size_t>(n_aux, (size_t(1) << 26) / per));
    double* bp = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_BPAD, pchunk * n_ao * ld * sizeof(double)));
    double* wt = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_WT, pchunk * n_out * ld * sizeof(double)));
    hipLaunchKernelGGL(rr::k_zero2, dim3(rr::grid_for(pchunk * n_out * ld)), dim3(256), 0, c.stream, wt,
                       pchunk * n_out * ld);
    for (size_t p0 = 0; p0 < n_aux; p0 += pchunk) {
      const size_t pc = std::min(pchunk, n_aux - p0);
      // B_P rows padded: bp[(P, m)][a] = in[P][m][a]
      hipLaunchKernelGGL(rr::k_pack2, dim3(rr::grid_for(pc * n_ao * ld)), dim3(256), 0, c.stream,
                         din + p0 * n_ao * n_ao, n_ao, 0, pc * n_ao, n_ao, ld, pc * n_ao, bp);
      RR_HIP(hipGetLastError());
      {  // W_P[m][o] = sum_a B_P[m][a] T[a][o], stored as wt_P[o][m]
        rr::GemmArgs g;
        g.A = bp, g.lda = (long long)ld, g.sAb = (long long)(n_ao * ld);
        g.B = tt, g.ldb = (long long)ld, g.sBb = 0;
        g.C = wt, g.ldcm = 1, g.ldcn = (long long)ld, g.sCb = (long long)(n_out * ld);
        g.M = (int)n_ao, g.N = (int)n_out, g.K = (int)ld, g.nbatch = (int)pc;
        rr::GemmSpec sp;
        sp.cfg = 1;
        rr::gemm_auto(c, sp, g);
      }
      {  // out_P[o][o'] = sum_m T[m][o] W_P[m][o']
        rr::GemmArgs g;
        g.A = tt, g.lda = (long long)ld, g.sAb = 0;
        g.B = wt, g.ldb = (long long)ld, g.sBb = (long long)(n_out * ld);
        g.C = ob.dev + p0 * n_out * n_out, g.ldcm = (long long)n_out, g.ldcn = 1, g.sCb = (long long)(n_out * n_out);
        g.M = (int)n_out, g.N = (int)n_out, g.K = (int)ld, g.nbatch = (int)pc;
        rr::GemmSpec sp;
        sp.cfg = 1;
        rr::gemm_auto(c, sp, g);
      }
    }
    rr::finish_out(c, ob);
  });
}

int rr_canonical_orthogonalize(rr_handle h, size_t n, const double* s, double threshold, double* x, size_t* n_kept) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(s && x && n_kept, RR_ERR_ARG, "canonicalOrthogonalize: null argument");
    *n_kept = 0;
    if (n == 0) return;
    double* a = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_S, n * n * sizeof(double)));
    RR_HIP(hipMemcpyAsync(a, s, n * n * sizeof(double), hipMemcpyDefault, c.stream));
    double* w = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_W, n * sizeof(double)));
    rr::eig_device(c, false, n, a, w);  // row j of a = eigenvector j, ascending w
    std::vector<double> wh(n);
    RR_HIP(hipMemcpyAsync(wh.data(), w, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
    size_t first = 0;
    while (first < n && !(wh[first] > threshold)) ++first;
    const size_t neff = n - first;
    *n_kept = neff;
    rr::OutBuf ob = rr::out_buf(c, rr::WS_TR_OUT, x, n * std::max<size_t>(neff, 1));
    if (neff == 0) {
      rr::finish_out(c, ob);
      return;
    }
    if (neff == n) {
      // Nothing dropped: the symmetric S^-1/2 = sum_k v_k v_k^T / sqrt(w_k) (LinearAlgebra.cpp:296-315).
      const size_t ld = rr::roundup(n, 8);
      double* y = static_cast<double*>(rr::wsbuf(c, rr::WS_TR_XT, n * ld * sizeof(double)));
      hipLaunchKernelGGL(rr::k_quarter, dim3(rr::grid_for(n * ld)), dim3(256), 0, c.stream, a, w, n, ld, y);
      RR_HIP(hipGetLastError());
      rr::GemmArgs g;
      g.A = y, g.lda = (long long)ld;
      g.B = y, g.ldb = (long long)ld;
      g.C = ob.dev, g.ldcm = (long long)n, g.ldcn = 1;
      g.M = (int)n, g.N = (int)n, g.K = (int)ld;
      rr::GemmSpec sp;
      sp.cfg = 1;
      rr::run_gemm(c, sp, g);
    } else {
      hipLaunchKernelGGL(rr::k_canon_x, dim3(rr::grid_for(n * neff)), dim3(256), 0, c.stream, a, w, n, first, neff, ob.dev);
      RR_HIP(hipGetLastError());
    }
    rr::finish_out(c, ob);
  });
}

}  // extern "C"
