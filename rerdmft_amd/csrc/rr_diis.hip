// librerdmft_gpu, the SCF loops' DIIS on the device (src/Utils/DIIS.h:13-46, DIIS.cpp): the
// commutator error e = F (P S) - (S P) F and Pulay's extrapolation with the (error, Fock) history
// kept in HBM. Only the (m+1) x (m+1) DIIS system is solved on the host; the inner products
// B_ij = Re <e_i|e_j> are reduced on the device (the new row only, the rest is kept from earlier
// calls) and the extrapolated matrix sum_i w_i F_i is formed on the device.
#include <algorithm>
#include <cmath>
#include <deque>
#include <limits>

#include "rr_core.hpp"

struct rr_diis {
  size_t max_vectors = 0;
  size_t n_err = 0, n_val = 0;    // doubles per stored error / value (a complex value counts twice)
  std::deque<double*> errors;     // newest first, device arrays
  std::deque<double*> values;
  std::vector<std::vector<double>> b;  // B_ij for the stored history, same order (newest first)
  double* dots = nullptr;         // device partial sums of the new row of B
  std::vector<double> last_weights;
};

namespace {

using rr::Context;
using rr::Error;

enum DiisWs { WS_DI_FPS = rr::WS_SIGMA_BASE + 5000, WS_DI_SPF, WS_DI_T, WS_DI_ERR, WS_DI_VAL };

constexpr int kDotBlocks = 128;

// part[j * kDotBlocks + b] = block b's share of sum_k a[k] * e_j[k] for the n stored errors e_j
// (real parts of conj(a) e_j for interleaved complex data: the plain product of the double arrays).
__global__ __launch_bounds__(256) void k_diis_dots(const double* a, const double* const* es, int n, size_t count,
                                                   double* part) {
  __shared__ double sh[256];
  for (int j = 0; j < n; ++j) {
    const double* e = es[j];
    double s = 0.0;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x)
      s += a[k] * e[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[j * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

struct Combo {
  const double* v[16];
  double w[16];
  int n;
};
// out = sum_i w_i v_i
__global__ void k_diis_combine(Combo c, size_t count, double* out) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < c.n; ++i) s += c.w[i] * c.v[i][k];
    out[k] = s;
  }
}

__global__ void k_diis_sub(const double* a, const double* b, size_t count, double* out) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x)
    out[k] = a[k] - b[k];
}

// Solves the (m+1) x (m+1) Pulay system by an LU factorisation with row pivoting (P A = L U, the unit
// lower factor kept in place of the eliminated entries, rows addressed through a permutation table),
// then forward and back substitution. The system is refused, as in the reference (DIIS.cpp:18-45),
// when a pivot is exactly zero or when min_k |U_kk| is not above tolerance * max_k |U_kk|; the pivot
// of a column is the first remaining row of largest magnitude, so the U diagonal, and with it that
// decision, is the one partial pivoting defines.
bool lu_solve_pivot_ratio(std::vector<double> a, const std::vector<double>& b, size_t n, double tolerance,
                          std::vector<double>& x) {
  std::vector<size_t> row(n);
  for (size_t i = 0; i < n; ++i) row[i] = i;
  double umax = 0.0, umin = std::numeric_limits<double>::infinity();
  for (size_t c = 0; c < n; ++c) {
    size_t p = c;
    double big = std::fabs(a[row[c] * n + c]);
    for (size_t r = c + 1; r < n; ++r) {
      const double mag = std::fabs(a[row[r] * n + c]);
      if (mag > big) {
        big = mag;
        p = r;
      }
    }
    std::swap(row[c], row[p]);
    if (big == 0.0) return false;
    umax = std::max(umax, big);
    umin = std::min(umin, big);
    const double* urow = a.data() + row[c] * n;
    for (size_t r = c + 1; r < n; ++r) {
      double* lrow = a.data() + row[r] * n;
      const double l = lrow[c] / urow[c];
      lrow[c] = l;
      for (size_t j = c + 1; j < n; ++j) lrow[j] -= l * urow[j];
    }
  }
  if (!(umin > tolerance * umax)) return false;
  std::vector<double> y(n);
  for (size_t i = 0; i < n; ++i) {
    const double* lrow = a.data() + row[i] * n;
    double s = b[row[i]];
    for (size_t k = 0; k < i; ++k) s -= lrow[k] * y[k];
    y[i] = s;
  }
  x.assign(n, 0.0);
  for (size_t i = n; i-- > 0;) {
    const double* urow = a.data() + row[i] * n;
    double s = y[i];
    for (size_t j = i + 1; j < n; ++j) s -= urow[j] * x[j];
    x[i] = s / urow[i];
  }
  return true;
}

void drop_history(rr_diis* d) {
  for (double* p : d->errors) (void)hipFree(p);
  for (double* p : d->values) (void)hipFree(p);
  d->errors.clear();
  d->values.clear();
  d->b.clear();
  d->last_weights.clear();
}

}  // namespace

using rr::guarded;

extern "C" {

int rr_scf_commutator_error(rr_handle h, int is_complex, size_t n, const double* fock, const double* density,
                            const double* overlap, double* error) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(fock && density && overlap && error, RR_ERR_ARG, "scfCommutatorError: null input");
    const int cw = is_complex ? 2 : 1;
    const size_t cnt = n * n * cw;
    if (n == 0) return;
    const double* dF = rr::in_buf(c, rr::WS_H, fock, cnt);
    const double* dP = rr::in_buf(c, rr::WS_P, density, cnt);
    const double* dS = rr::in_buf(c, rr::WS_S, overlap, cnt);
    double* t = static_cast<double*>(rr::wsbuf(c, WS_DI_T, cnt * sizeof(double)));
    double* fps = static_cast<double*>(rr::wsbuf(c, WS_DI_FPS, cnt * sizeof(double)));
    double* spf = static_cast<double*>(rr::wsbuf(c, WS_DI_SPF, cnt * sizeof(double)));
    const bool cx = is_complex != 0;
    rr::dgemm_small(c, cx, 'N', 'N', n, n, n, dP, n, dS, n, t, n);    // P S
    rr::dgemm_small(c, cx, 'N', 'N', n, n, n, dF, n, t, n, fps, n);   // F (P S)
    rr::dgemm_small(c, cx, 'N', 'N', n, n, n, dS, n, dP, n, t, n);    // S P
    rr::dgemm_small(c, cx, 'N', 'N', n, n, n, t, n, dF, n, spf, n);   // (S P) F
    rr::OutBuf oe = rr::out_buf(c, WS_DI_ERR, error, cnt);
    hipLaunchKernelGGL(k_diis_sub, dim3(rr::grid_for(cnt)), dim3(256), 0, c.stream, fps, spf, cnt, oe.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, oe);
  });
}

int rr_diis_create(rr_handle h, size_t max_vectors, rr_diis_t* out) {
  if (!out) return RR_ERR_ARG;
  *out = nullptr;
  rr_diis* d = new rr_diis();
  const int rc = guarded(h, [&](Context&) {
    RR_REQUIRE(max_vectors <= 16, RR_ERR_ARG, "Diis: at most 16 stored vectors");
    d->max_vectors = max_vectors;
    RR_HIP(hipMalloc(&d->dots, 16 * kDotBlocks * sizeof(double)));
  });
  if (rc != RR_OK) {
    if (d->dots) (void)hipFree(d->dots);
    delete d;
    return rc;
  }
  *out = d;
  return RR_OK;
}

int rr_diis_destroy(rr_handle h, rr_diis_t d) {
  if (!h || !d) return RR_ERR_ARG;
  return guarded(h, [&](Context& c) {
    RR_HIP(hipStreamSynchronize(c.stream));
    drop_history(d);
    if (d->dots) (void)hipFree(d->dots);
    delete d;
  });
}

int rr_diis_clear(rr_handle h, rr_diis_t d) {
  if (!h || !d) return RR_ERR_ARG;
  return guarded(h, [&](Context& c) {
    RR_HIP(hipStreamSynchronize(c.stream));
    drop_history(d);
  });
}

int rr_diis_extrapolate(rr_handle h, rr_diis_t d, size_t n_error, const double* error, size_t n_value,
                        const double* value, double* out, double* weights, size_t* n_weights) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(d && error && value && out, RR_ERR_ARG, "Diis::extrapolate: null input");
    if (n_weights) *n_weights = 0;
    d->last_weights.clear();
    const double* dv = rr::in_buf(c, WS_DI_VAL, value, n_value);
    rr::OutBuf ov = rr::out_buf(c, rr::WS_OUT_F, out, n_value);
    auto pass_through = [&]() {
      RR_HIP(hipMemcpyAsync(ov.dev, dv, n_value * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
      rr::finish_out(c, ov);
    };
    if (d->max_vectors < 2) {
      pass_through();
      return;
    }
    if (!d->errors.empty())
      RR_REQUIRE(d->n_err == n_error && d->n_val == n_value, RR_ERR_DIM,
                 "Diis::extrapolate: shapes differ from the stored history");
    d->n_err = n_error, d->n_val = n_value;
    // store the new pair (newest first), dropping the oldest beyond max_vectors
    double* e_new = nullptr;
    double* v_new = nullptr;
    if (d->errors.size() == d->max_vectors) {
      e_new = d->errors.back(), v_new = d->values.back();
      d->errors.pop_back(), d->values.pop_back(), d->b.pop_back();
      for (auto& row : d->b) row.pop_back();
    } else {
      RR_HIP(hipMalloc(&e_new, std::max<size_t>(n_error, 1) * sizeof(double)));
      RR_HIP(hipMalloc(&v_new, std::max<size_t>(n_value, 1) * sizeof(double)));
    }
    RR_HIP(hipMemcpyAsync(e_new, error, n_error * sizeof(double), hipMemcpyDefault, c.stream));
    RR_HIP(hipMemcpyAsync(v_new, dv, n_value * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    d->errors.push_front(e_new), d->values.push_front(v_new);
    // new row of B: Re <e_new|e_j> for every stored j (including itself)
    const size_t total = d->errors.size();
    const double** ptrs = static_cast<const double**>(rr::wsbuf(c, WS_DI_ERR + 100, 16 * sizeof(double*)));
    std::vector<const double*> hp(d->errors.begin(), d->errors.end());
    RR_HIP(hipMemcpyAsync(ptrs, hp.data(), total * sizeof(double*), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_diis_dots, dim3(kDotBlocks), dim3(256), 0, c.stream, e_new, ptrs, (int)total, n_error, d->dots);
    RR_HIP(hipGetLastError());
    std::vector<double> part(total * kDotBlocks);
    RR_HIP(hipMemcpyAsync(part.data(), d->dots, part.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
    std::vector<double> row(total, 0.0);
    for (size_t j = 0; j < total; ++j)
      for (int b = 0; b < kDotBlocks; ++b) row[j] += part[j * kDotBlocks + b];
    for (size_t i = 1; i < total; ++i) d->b[i - 1].insert(d->b[i - 1].begin(), row[i]);
    d->b.insert(d->b.begin(), row);
    // use the newest m pairs; drop the oldest whenever the system is singular (DIIS.cpp:110-135)
    for (size_t m = total; m >= 2; --m) {
      double scale = 0.0;
      for (size_t i = 0; i < m; ++i) scale = std::max(scale, d->b[i][i]);
      if (!(scale > 0.0)) break;  // zero error: nothing to extrapolate
      const size_t n = m + 1;
      std::vector<double> a(n * n, 0.0), rhs(n, 0.0), x;
      for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < m; ++j) a[i * n + j] = d->b[i][j] / scale;
      for (size_t i = 0; i < m; ++i) a[i * n + m] = a[m * n + i] = -1.0;
      rhs[m] = -1.0;
      if (!lu_solve_pivot_ratio(std::move(a), std::move(rhs), n, 1e-14, x)) continue;
      Combo cb;
      cb.n = (int)m;
      for (size_t i = 0; i < m; ++i) cb.v[i] = d->values[i], cb.w[i] = x[i];
      hipLaunchKernelGGL(k_diis_combine, dim3(rr::grid_for(n_value)), dim3(256), 0, c.stream, cb, n_value, ov.dev);
      RR_HIP(hipGetLastError());
      rr::finish_out(c, ov);
      d->last_weights.assign(x.begin(), x.begin() + (std::ptrdiff_t)m);
      if (weights) std::copy(d->last_weights.begin(), d->last_weights.end(), weights);
      if (n_weights) *n_weights = m;
      return;
    }
    pass_through();
  });
}

}  // extern "C"
