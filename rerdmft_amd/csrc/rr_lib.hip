// librerdmft_gpu: MI355X (gfx950) RI / Cholesky J/K Fock builds and RI AO->MO transforms behind a
// C ABI (include/rerdmft_gpu.h).
//
// Device layout of a 3-index tensor slot (an RI tensor (n_aux, n*n) or a set of Cholesky vectors):
//   Bt[P][nu][mu] = B[P][mu][nu]   (each aux row transposed), leading dimension ld = tensor_ld(n),
//   columns [n, ld) zero-filled, kept in any of fp64 / fp32 / bf16 (all copies resident in HBM).
// With this layout every contraction of the path is the "NT" form C = A B^T of rr_gemm.hpp:
//   exchange half-transform  W[nu][(P,i)] = sum_mu Bt_P[nu][mu] Lt[i][mu]      (A = Bt_P, B = Lt)
//   exchange build           K[mu][nu]    = sum_(P,i) W[mu][(P,i)] W[nu][(P,i)] (Hermitian rank-k)
//   Coulomb                  gamma_P = <Bt_P, P^T>,  Jt = sum_P gamma_P Bt_P      (HBM streams)
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "rr_core.hpp"
#include "rr_gemm_rows.hpp"

namespace rr {
// Named grow-only device workspace buffers.
// Scratch region of the allocation-free entry point (defined with it below).
hipError_t register_scratch(int device);

void* wsbuf(Context& c, int id, size_t bytes) {
  if (c.arena) {
    auto at = c.arena_ws.find(id);
    if (at != c.arena_ws.end() && at->second.second >= bytes) return at->second.first;
    const size_t off = roundup(c.arena_used, 256), sz = std::max<size_t>(bytes, 256);
    if (off + sz > c.arena_size) {
      char msg[200];
      snprintf(msg, sizeof(msg), "allocation-free call: the workspace needs more than the %zu MiB scratch region",
               c.arena_size >> 20);
      throw Error{RR_ERR_ARG, msg};
    }
    c.arena_used = off + sz;
    c.arena_ws[id] = {c.arena + off, sz};
    return c.arena + off;
  }
  auto it = c.ws.find(id);
  if (it != c.ws.end() && it->second.second >= bytes) return it->second.first;
  if (it != c.ws.end()) {
    RR_HIP(hipStreamSynchronize(c.stream));
    RR_HIP(hipFree(it->second.first));
    c.ws.erase(it);
  }
  void* p = nullptr;
  RR_HIP(hipMalloc(&p, std::max<size_t>(bytes, 256)));
  c.ws[id] = {p, std::max<size_t>(bytes, 256)};
  ++c.ws_generation;  // captured graphs that used the old buffers must be rebuilt
  return p;
}


// True when p is device (or managed) memory visible to the current device.
bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}


// ================================================================================================
// Elementwise / layout kernels
// ================================================================================================
template <typename T>
__device__ inline T from_double(double v) {
  return Conv<T>::to(v);
}

// Transposed store of `rows` aux rows: out[P][nu][mu] = in[P][mu][nu], zero-fill columns [n, ld).
template <typename TO>
__global__ void k_store_rows(const double* in, size_t n, size_t ld, TO* out) {
  __shared__ double tile[32][33];
  const size_t P = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const size_t mu0 = blockIdx.x * 32, nu0 = blockIdx.y * 32;
  const double* src = in + P * n * n;
  for (int yy = ty; yy < 32; yy += 8) {
    const size_t mu = mu0 + yy, nu = nu0 + tx;
    tile[yy][tx] = (mu < n && nu < n) ? src[mu * n + nu] : 0.0;
  }
  __syncthreads();
  TO* dst = out + P * n * ld;
  for (int yy = ty; yy < 32; yy += 8) {
    const size_t nu = nu0 + yy, mu = mu0 + tx;
    if (nu < n && mu < ld) dst[nu * ld + mu] = from_double<TO>(tile[tx][yy]);
  }
}

__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Synthetic symmetric 3-index tensor: value(P, mu, nu) = sigma * N(0,1) * exp(-|mu-nu| / decay),
// N(0,1) by Box-Muller from a counter-based hash of (seed, P, min(mu,nu), max(mu,nu)).
template <typename TO>
__global__ void k_generate(size_t naux, size_t n, size_t ld, double sigma, double decay, uint64_t seed, size_t p0,
                           size_t prows, TO* out) {
  const size_t total = prows * n * ld;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t P = p0 + idx / (n * ld);
    const size_t rem = idx % (n * ld);
    const size_t nu = rem / ld, mu = rem % ld;
    double v = 0.0;
    if (mu < n) {
      const size_t a = mu < nu ? mu : nu, b = mu < nu ? nu : mu;
      const uint64_t key = splitmix64(seed ^ splitmix64(((uint64_t)P * n + a) * n + b));
      const uint64_t k2 = splitmix64(key);
      const double u1 = ((key >> 11) + 1) * (1.0 / 9007199254740993.0);
      const double u2 = (k2 >> 11) * (1.0 / 9007199254740992.0);
      const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      const double d = (double)(b - a);
      v = sigma * z * exp(-d / decay);
    }
    out[(P - p0) * n * ld + rem] = from_double<TO>(v);
  }
}

template <typename TI, typename TO>
__global__ void k_convert(const TI* in, TO* out, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    out[i] = from_double<TO>(Conv<TI>::from(in[i]));
}

// out[r][c] (ld_out) = scale * op(in)[r][c] for r < Rv, c < Cc (op = transpose when trans); rows
// [Rv, R) and columns [Cc, ld_out) are zero-filled.
template <typename TO>
__device__ __forceinline__ void k_pack_body(const double* in, size_t ld_in, int trans, size_t R, size_t Rv, size_t Cc, size_t ld_out,
                       double scale, TO* out, unsigned bx_, unsigned by_, unsigned gx_) {
  const size_t total = R * ld_out;
  for (size_t idx = bx_ * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gx_ * blockDim.x) {
    const size_t r = idx / ld_out, c = idx % ld_out;
    double v = 0.0;
    if (c < Cc && r < Rv) v = scale * (trans ? in[c * ld_in + r] : in[r * ld_in + c]);
    out[idx] = from_double<TO>(v);
  }
}
template <typename TO>
__global__ void k_pack(const double* in, size_t ld_in, int trans, size_t R, size_t Rv, size_t Cc, size_t ld_out,
                       double scale, TO* out) {
  k_pack_body<TO>(in, ld_in, trans, R, Rv, Cc, ld_out, scale, out, blockIdx.x, blockIdx.y, gridDim.x);
}

__global__ void k_conj_inplace(double* a, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    a[2 * i + 1] = -a[2 * i + 1];
}

__global__ void k_zero(double* p, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}


// ================================================================================================
// Coulomb streams
// ================================================================================================
template <typename TB>
struct Vec16;  // 16-byte vector load, widened to double
template <>
struct Vec16<double> {
  static constexpr int N = 2;
  __device__ static void load(const double* p, double* o) {
    d2 v = *reinterpret_cast<const d2*>(p);
    o[0] = v.x, o[1] = v.y;
  }
  // streaming form for data read once per build (the 3-index tensor): non-temporal loads keep it
  // from displacing reused lines and stream about 12% faster from HBM
  __device__ static void load_nt(const double* p, double* o) { widen(raw_nt(p), o); }
  // raw 16-byte vector (widened at use, so a thread keeps many loads in flight in few registers)
  typedef d2 raw_t;
  __device__ static raw_t raw_nt(const double* p) { return __builtin_nontemporal_load(reinterpret_cast<const d2*>(p)); }
  __device__ static raw_t raw(const double* p) { return *reinterpret_cast<const d2*>(p); }
  __device__ static void widen(raw_t v, double* o) { o[0] = v.x, o[1] = v.y; }
};
template <>
struct Vec16<float> {
  static constexpr int N = 4;
  __device__ static void load(const float* p, double* o) {
    f4 v = *reinterpret_cast<const f4*>(p);
    o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w;
  }
  __device__ static void load_nt(const float* p, double* o) { widen(raw_nt(p), o); }
  typedef f4 raw_t;
  __device__ static raw_t raw_nt(const float* p) { return __builtin_nontemporal_load(reinterpret_cast<const f4*>(p)); }
  __device__ static raw_t raw(const float* p) { return *reinterpret_cast<const f4*>(p); }
  __device__ static void widen(raw_t v, double* o) { o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w; }
};
template <>
struct Vec16<bf16_t> {
  static constexpr int N = 8;
  __device__ static void load(const bf16_t* p, double* o) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    widen(*reinterpret_cast<const u4*>(p), o);
  }
  __device__ static void load_nt(const bf16_t* p, double* o) { widen(raw_nt(p), o); }
  typedef unsigned int raw_t __attribute__((ext_vector_type(4)));
  __device__ static raw_t raw_nt(const bf16_t* p) { return __builtin_nontemporal_load(reinterpret_cast<const raw_t*>(p)); }
  __device__ static raw_t raw(const bf16_t* p) { return *reinterpret_cast<const raw_t*>(p); }
  __device__ static void widen(raw_t v, double* o) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      union {
        unsigned u;
        float f;
      } lo, hi;
      lo.u = v[i] << 16;
      hi.u = v[i] & 0xffff0000u;
      o[2 * i] = lo.f, o[2 * i + 1] = hi.f;
    }
  }
};

__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
  return s;
}

// gpart[split][P] = sum_{e in split's chunk} Bt[P][e] * Dt[e], R aux rows per workgroup.
template <typename TB, int R>
__global__ __launch_bounds__(256) void k_gamma(const TB* B, size_t rowlen, size_t naux, const TB* Dt, size_t chunk,
                                               double* gpart) {
  using V = Vec16<TB>;
  constexpr int VN = V::N;
  __shared__ double sh[8];
  const size_t P0 = (size_t)blockIdx.x * R;
  const size_t split = blockIdx.y;
  const size_t e0 = split * chunk;
  const size_t e1 = std::min(rowlen, e0 + chunk);
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  for (size_t e = e0 + (size_t)threadIdx.x * VN; e < e1; e += 256 * VN) {
    double dv[VN];
    V::load(Dt + e, dv);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (P0 + r < naux) {
        double bv[VN];
        V::load_nt(B + (P0 + r) * rowlen + e, bv);
#pragma unroll
        for (int q = 0; q < VN; ++q) acc[r] += bv[q] * dv[q];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double s = block_sum(acc[r], sh);
    if (threadIdx.x == 0 && P0 + r < naux) gpart[split * naux + P0 + r] = s;
  }
}

// gamma[P] = sum_split gpart[split][P] (+ gamma[P] when accumulate)
__global__ void k_gamma_reduce(const double* gpart, int nsplit, size_t naux, double* gamma, int accumulate) {
  for (size_t P = blockIdx.x * (size_t)blockDim.x + threadIdx.x; P < naux; P += (size_t)gridDim.x * blockDim.x) {
    double s = accumulate ? gamma[P] : 0.0;
    for (int z = 0; z < nsplit; ++z) s += gpart[(size_t)z * naux + P];
    gamma[P] = s;
  }
}

// jpart[split][e] = sum_{P in split range} gamma[P] * Bt[P][e]; each thread owns U 16-byte vectors.
template <typename TB, int U, int PU = (sizeof(TB) == 8 ? 2 : 1)>
__device__ __forceinline__ void k_coulomb_body(const TB* B, size_t rowlen, size_t naux, const double* gamma,
                                                 size_t p_per_split, double* jpart, unsigned bx_, unsigned by_, unsigned gx_) {
  using V = Vec16<TB>;
  constexpr int VN = V::N;
  const size_t base = (size_t)bx_ * (256 * VN * U) + (size_t)threadIdx.x * VN;
  const size_t split = by_;
  const size_t pb = split * p_per_split, pe = std::min(naux, pb + p_per_split);
  double acc[U][VN];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int q = 0; q < VN; ++q) acc[u][q] = 0.0;
  bool valid[U];
#pragma unroll
  for (int u = 0; u < U; ++u) valid[u] = base + (size_t)u * 256 * VN < rowlen;
  // FP64: two aux rows per iteration, both rows' loads issued before either row is accumulated (the
  // sum still runs over P in order); narrower types keep one row (the wider rows cost occupancy)
  size_t P = pb;
  if constexpr (PU == 2)
  for (; P + 1 < pe; P += 2) {
    double bv[2][U][VN];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (valid[u]) V::load_nt(B + (P + h) * rowlen + base + (size_t)u * 256 * VN, bv[h][u]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double g = gamma[P + h];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (valid[u]) {
#pragma unroll
          for (int q = 0; q < VN; ++q) acc[u][q] += g * bv[h][u][q];
        }
    }
  }
  for (; P < pe; ++P) {
    const double g = gamma[P];
    const TB* row = B + P * rowlen;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (valid[u]) {
        double bv[VN];
        V::load_nt(row + base + (size_t)u * 256 * VN, bv);
#pragma unroll
        for (int q = 0; q < VN; ++q) acc[u][q] += g * bv[q];
      }
    }
  }
  double* out = jpart + split * rowlen;
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (valid[u]) {
#pragma unroll
      for (int q = 0; q < VN; ++q) out[base + (size_t)u * 256 * VN + q] = acc[u][q];
    }
}
template <typename TB, int U>
__global__ __launch_bounds__(256) void k_coulomb(const TB* B, size_t rowlen, size_t naux, const double* gamma,
                                                 size_t p_per_split, double* jpart) {
  k_coulomb_body<TB, U>(B, rowlen, naux, gamma, p_per_split, jpart, blockIdx.x, blockIdx.y, gridDim.x);
}

// Upper triangle of a symmetric tensor row in 16-byte vectors: row nu holds the vectors from the one
// containing the diagonal (index floor(nu / VN)) to the end of the padded row. Rows come in groups of
// VN with W - g vectors each (group g, W = ld / VN), so the vectors before group g number
// S(g) = VN (g W - g (g - 1) / 2); tri_offset maps a flat vector index to its element offset.
__device__ inline size_t tri_offset(size_t v, size_t W, size_t VN, size_t ld) {
  const double b = 2.0 * static_cast<double>(W) + 1.0;
  const double disc = b * b - 8.0 * static_cast<double>(v) / static_cast<double>(VN);
  size_t g = static_cast<size_t>((b - sqrt(disc > 0.0 ? disc : 0.0)) * 0.5);
  auto S = [&](size_t x) { return VN * (x * W) - VN * ((x * (x - 1)) / 2); };
  while (g > 0 && S(g) > v) --g;
  while (g + 1 <= W && S(g + 1) <= v) ++g;
  const size_t off = v - S(g), len = W - g;
  return (g * VN + off / len) * ld + (g + off % len) * VN;
}
// Vectors of the upper triangle of an n x ld row.
inline size_t tri_vectors(size_t n, size_t ld, size_t VN) {
  const size_t W = ld / VN;
  size_t t = 0;
  for (size_t nu = 0; nu < n; ++nu) t += W - nu / VN;
  return t;
}

// k_coulomb over the upper triangle only (symmetric tensors): jpart[split][e] for the triangle
// elements e; the rest of jpart is left unwritten.
template <typename TB, int U, int PU = (sizeof(TB) == 8 ? 2 : 1)>
__device__ __forceinline__ void k_coulomb_tri_body(const TB* B, size_t rowlen, size_t ld, size_t nvec, size_t naux,
                                                     const double* gamma, size_t p_per_split, double* jpart, unsigned bx_, unsigned by_, unsigned gx_) {
  using V = Vec16<TB>;
  constexpr int VN = V::N;
  const size_t W = ld / VN;
  const size_t vbase = (size_t)bx_ * (256 * U) + threadIdx.x;
  const size_t split = by_;
  const size_t pb = split * p_per_split, pe = std::min(naux, pb + p_per_split);
  size_t e[U];
  bool valid[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t v = vbase + (size_t)u * 256;
    valid[u] = v < nvec;
    e[u] = valid[u] ? tri_offset(v, W, VN, ld) : 0;
  }
  double acc[U][VN];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int q = 0; q < VN; ++q) acc[u][q] = 0.0;
  size_t P = pb;  // PU aux rows per iteration: PU * U loads in flight per thread (sums still in P order)
  if constexpr (PU > 1)
  for (; P + PU <= pe; P += PU) {
    typename V::raw_t bv[PU][U];
#pragma unroll
    for (int h = 0; h < PU; ++h)
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (valid[u]) bv[h][u] = V::raw_nt(B + (P + h) * rowlen + e[u]);
#pragma unroll
    for (int h = 0; h < PU; ++h) {
      const double g = gamma[P + h];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (valid[u]) {
          double w[VN];
          V::widen(bv[h][u], w);
#pragma unroll
          for (int q = 0; q < VN; ++q) acc[u][q] += g * w[q];
        }
    }
  }
  for (; P < pe; ++P) {
    const double g = gamma[P];
    const TB* row = B + P * rowlen;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (valid[u]) {
        double bv[VN];
        V::load_nt(row + e[u], bv);
#pragma unroll
        for (int q = 0; q < VN; ++q) acc[u][q] += g * bv[q];
      }
    }
  }
  double* out = jpart + split * rowlen;
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (valid[u]) {
#pragma unroll
      for (int q = 0; q < VN; ++q) out[e[u] + q] = acc[u][q];
    }
}
template <typename TB, int U, int PU>
__global__ __launch_bounds__(256) void k_coulomb_tri(const TB* B, size_t rowlen, size_t ld, size_t nvec, size_t naux,
                                                     const double* gamma, size_t p_per_split, double* jpart) {
  k_coulomb_tri_body<TB, U, PU>(B, rowlen, ld, nvec, naux, gamma, p_per_split, jpart, blockIdx.x, blockIdx.y, gridDim.x);
}
// Upper-triangle stream shape per operand type (vectors per thread, aux rows per iteration), chosen by
// measurement at n = 1000, n_aux = 4000: fewer vectors and more rows per thread keep more loads in
// flight at a higher occupancy (bf16 5.7 TB/s, fp64 6.0 TB/s standalone).
template <typename TB>
struct CoulTri;
template <>
struct CoulTri<double> {
  static constexpr int U = 2, PU = 8;
};
template <>
struct CoulTri<float> {
  static constexpr int U = 2, PU = 4;
};
template <>
struct CoulTri<bf16_t> {
  static constexpr int U = 2, PU = 2;
};

// ================================================================================================
// Epilogues
// ================================================================================================
// Real NON_REL assembly: J[mu][nu] = sum_s Jt_s[nu*ld + mu]; K from split-K partials of a
// lower-triangle-tiled (tri) or full (sym_sum: K = G + G^T) rank-k update; F = h + J + kfac*K.
struct RealAsm {
  size_t n = 0, ld = 0;
  const double* jpart = nullptr;
  int jsplit = 0;
  size_t jstride = 0;
  const double* kpart = nullptr;
  int ksplit = 0;
  size_t kstride = 0;
  int ktile = 0;   // tile size of the tri layout (0: full G)
  int ksym = 0;    // full G only: 1 -> K = G + G^T, 0 -> K = G
  int jtri = 0;    // jpart holds the upper triangle only (row <= column)
  double kfac = -0.5;
  const double* h = nullptr;
  double* F = nullptr;
  double* J = nullptr;
  double* K = nullptr;
};

__device__ __forceinline__ void k_real_asm_body(RealAsm a, unsigned bx_, unsigned by_, unsigned gx_) {
  const size_t total = a.n * a.n;
  for (size_t idx = bx_ * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gx_ * blockDim.x) {
    const size_t mu = idx / a.n, nu = idx % a.n;
    double j = 0.0;
    if (a.jpart) {
      const size_t src = a.jtri ? (mu < nu ? mu * a.ld + nu : nu * a.ld + mu) : nu * a.ld + mu;
      for (int s = 0; s < a.jsplit; ++s) j += a.jpart[s * a.jstride + src];
    }
    double k = 0.0;
    if (a.kpart) {
      if (a.ktile > 0) {
        // lower-triangle tiling: every element with mu >= nu was computed, for any square tile size
        const bool lower = mu >= nu;
        const size_t src = lower ? mu * a.n + nu : nu * a.n + mu;
        for (int s = 0; s < a.ksplit; ++s) k += a.kpart[s * a.kstride + src];
      } else if (a.ksym) {
        for (int s = 0; s < a.ksplit; ++s) k += a.kpart[s * a.kstride + mu * a.n + nu] + a.kpart[s * a.kstride + nu * a.n + mu];
      } else {
        for (int s = 0; s < a.ksplit; ++s) k += a.kpart[s * a.kstride + mu * a.n + nu];
      }
    }
    if (a.J) a.J[idx] = j;
    if (a.K) a.K[idx] = k;
    if (a.F) a.F[idx] = (a.h ? a.h[idx] : 0.0) + j + a.kfac * k;
  }
}
__global__ void k_real_asm(RealAsm a) {
  k_real_asm_body(a, blockIdx.x, blockIdx.y, gridDim.x);
}

// Complex Hermitian assembly over nb diagonal Coulomb blocks: out[a][b] (complex, dim nd) =
//   h + J_blk (on the diagonal blocks listed in jblk) - K, K from tri-tiled complex partials.
struct CplxAsm {
  size_t nd = 0;
  const double* kpart = nullptr;  // complex interleaved, partials [ksplit][nd][nd]
  int ksplit = 0;
  size_t kstride = 0;             // in complex elements
  int ktile = 0;
  double kfac = -1.0;
  const double* h = nullptr;      // complex nd x nd (may be null)
  double* out = nullptr;          // complex nd x nd
  // Coulomb: up to 4 diagonal blocks; block b occupies [off[b], off[b]+sz[b]) and reads
  // J = jt[...] from jpart (real, transposed layout with leading dimension jld[b])
  int nblk = 0;
  size_t off[4] = {0, 0, 0, 0}, sz[4] = {0, 0, 0, 0}, jld[4] = {0, 0, 0, 0};
  const double* jsrc[4] = {nullptr, nullptr, nullptr, nullptr};  // reduced Jt for each block
};

__global__ void k_cplx_asm(CplxAsm a) {
  const size_t total = a.nd * a.nd;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t r = idx / a.nd, c = idx % a.nd;
    double kr = 0.0, ki = 0.0;
    if (a.kpart) {
      const bool lower = r >= c;  // Hermitian lower-triangle tiling (any square tile size)
      const size_t src = lower ? r * a.nd + c : c * a.nd + r;
      for (int s = 0; s < a.ksplit; ++s) {
        kr += a.kpart[2 * (s * a.kstride + src)];
        ki += a.kpart[2 * (s * a.kstride + src) + 1];
      }
      if (!lower) ki = -ki;
      if (r == c) ki = 0.0;  // a Hermitian K has a real diagonal (the product leaves rounding there)
    }
    double vr = a.kfac * kr, vi = a.kfac * ki;
    for (int b = 0; b < a.nblk; ++b) {
      if (r >= a.off[b] && r < a.off[b] + a.sz[b] && c >= a.off[b] && c < a.off[b] + a.sz[b]) {
        const size_t mu = r - a.off[b], nu = c - a.off[b];
        vr += a.jsrc[b][nu * a.jld[b] + mu];
      }
    }
    if (a.h) vr += a.h[2 * idx], vi += a.h[2 * idx + 1];
    a.out[2 * idx] = vr;
    a.out[2 * idx + 1] = vi;
  }
}

// Reduce split partials: out[e] = sum_s part[s*stride + e], summed in split order. Eight loads are
// issued before their adds so a thread has eight partials in flight.
__device__ __forceinline__ void k_reduce_parts_body(const double* part, int nsplit, size_t stride, size_t count, double* out, unsigned bx_, unsigned by_, unsigned gx_) {
  for (size_t i = bx_ * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gx_ * blockDim.x) {
    double s = 0.0;
    int z = 0;
    for (; z + 8 <= nsplit; z += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(z + u) * stride + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; z < nsplit; ++z) s += part[(size_t)z * stride + i];
    out[i] = s;
  }
}
__global__ void k_reduce_parts(const double* part, int nsplit, size_t stride, size_t count, double* out) {
  k_reduce_parts_body(part, nsplit, stride, count, out, blockIdx.x, blockIdx.y, gridDim.x);
}

// ================================================================================================
// GEMM dispatch (tile configurations instantiated for gfx950; the autotuner picks among them)
// ================================================================================================

struct TileCfg {
  int wm, wn;
};
// Index -> (WM, WN); each workgroup tile is (32*WM) x (32*WN).
static const TileCfg kTiles[] = {{1, 1}, {2, 2}, {4, 4}, {2, 4}, {4, 2}, {2, 7}, {1, 7}, {4, 1}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

template <typename TI, typename TO, bool CPLX>
void launch_gemm_cfg(int cfg, const GemmArgs& g, int grid, hipStream_t s) {
  switch (cfg) {
#define RR_CASE(i, WM, WN) \
  case i:                  \
    hipLaunchKernelGGL((gemm_nt_kernel<TI, TO, WM, WN, CPLX>), dim3(grid), dim3(256), 0, s, g); break;
    RR_CASE(0, 1, 1)
    RR_CASE(1, 2, 2)
    RR_CASE(2, 4, 4)
    RR_CASE(3, 2, 4)
    RR_CASE(4, 4, 2)
    RR_CASE(5, 2, 7)
    RR_CASE(6, 1, 7)
    RR_CASE(7, 4, 1)
#undef RR_CASE
    default:
      throw Error{RR_ERR_ARG, "bad tile configuration"};
  }
}


// LDS-staged kernel variants: (BM, BN, waves along M)
struct LdsCfg {
  int bm, bn, wgm;
  int bf16_only;  // single-buffered, higher-occupancy variants built for bf16 operands only
  int fp64_only;  // eight-wave fp64 / fp32 direct-to-LDS variants (not bf16)
  int rows = 0;   // persistent kernels of rr_gemm_rows.hpp: 1 row-merged half-transform (NT = bn / 16),
                  // 2 split-K rank-k update
};
static const LdsCfg kLds[] = {{128, 128, 2, 0}, {128, 208, 4, 0}, {256, 128, 2, 0}, {64, 208, 4, 0},
                              {64, 224, 2, 0},  {256, 208, 4, 0}, {128, 224, 2, 0}, {128, 208, 4, 1},
                              {128, 208, 4, 1}, {256, 208, 4, 1}, {128, 128, 2, 1}, {64, 208, 4, 1},
                              // multi-stage direct-to-LDS (bf16): 12-16
                              {128, 208, 4, 1}, {128, 208, 4, 1}, {64, 208, 4, 1}, {128, 128, 2, 1},
                              {256, 128, 2, 1},
                              // direct-to-LDS built with AGPR accumulators (rr_gemm_wide.hip): 17-19
                              {128, 416, 2, 1}, {256, 224, 2, 1}, {128, 208, 4, 1},
                              // lean direct-to-LDS (gemm_nt_dma2_kernel): 20-22 here, 23-24 AGPR form
                              {128, 208, 4, 1}, {256, 128, 2, 1}, {128, 128, 2, 1}, {128, 208, 4, 1},
                              {128, 416, 2, 1},
                              // lean direct-to-LDS, 2 x 2 waves of 64 x 112: 25
                              {128, 224, 2, 1},
                              // fp64 / fp32 direct-to-LDS with eight waves: 26-29
                              {128, 208, 8, 0, 1}, {128, 208, 8, 0, 1}, {128, 128, 4, 0, 1}, {128, 224, 4, 0, 1},
                              // lean bf16 direct-to-LDS with eight waves: 30-33
                              {256, 208, 8, 1}, {256, 208, 8, 1}, {256, 224, 4, 1}, {256, 224, 4, 1},
                              // bf16 through the eight-wave full-line kernel (two matrix-core steps per stage): 34-36
                              {256, 208, 8, 1}, {256, 224, 4, 1}, {128, 208, 8, 1},
                              // fp64 / fp32 eight-wave square tiles for the rank-k update: 37-38
                              {128, 128, 4, 0, 1}, {192, 192, 4, 0, 1},
                              // row-merged persistent bf16 half-transform, 128 x 400 / 416 / 448 tiles: 39-41
                              {128, 400, 2, 1, 0, 1}, {128, 416, 2, 1, 0, 1}, {128, 448, 2, 1, 0, 1},
                              // persistent split-K bf16 rank-k update, 256 x 256 tiles: 42
                              {256, 256, 2, 1, 0, 2}};
constexpr int kNumLds = sizeof(kLds) / sizeof(kLds[0]);

// rr_gemm_wide.hip: direct-to-LDS variants 17-19 and 23-24 compiled with AGPR accumulators.
bool launch_dma_agpr(int v, bool out_f64, const GemmArgs& g, int grid, hipStream_t s);

template <typename TI, typename TO>
void launch_lds(int v, const GemmArgs& g, int grid, hipStream_t s) {
  constexpr int SUB = 4 * Mma<TI>::KL;
  constexpr bool B16 = std::is_same<TI, bf16_t>::value;
  constexpr int BK2 = 2 * SUB;  // two fragment sub-steps per stage
  switch (v) {
    case 0: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 128, 128, (B16 ? BK2 : SUB), 2>), dim3(grid), dim3(256), 0, s, g); break;
    case 1: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 128, 208, SUB, 4>), dim3(grid), dim3(256), 0, s, g); break;
    case 3: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 64, 208, SUB, 4>), dim3(grid), dim3(256), 0, s, g); break;
    case 4: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 64, 224, (B16 ? BK2 : SUB), 2>), dim3(grid), dim3(256), 0, s, g); break;
    default:
      // the wide variants exceed the fp64 register budget and are only built for fp32 / bf16
      if constexpr (!std::is_same<TI, double>::value) {
        switch (v) {
          case 2: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 256, 128, (B16 ? BK2 : SUB), 2>), dim3(grid), dim3(256), 0, s, g); return;
          case 5: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 256, 208, SUB, 4>), dim3(grid), dim3(256), 0, s, g); return;
          case 6: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 128, 224, (B16 ? BK2 : SUB), 2>), dim3(grid), dim3(256), 0, s, g); return;
          default: break;
        }
      }
      if constexpr (std::is_same<TI, double>::value || std::is_same<TI, float>::value) {
        switch (v) {
          case 26: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 128, 208, 8, 2, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 27: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 128, 208, 8, 3, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 28: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 128, 128, 4, 3, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 29: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 128, 224, 4, 2, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 37: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 128, 128, 4, 4, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 38: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 192, 192, 4, 3, 8>), dim3(grid), dim3(512), 0, s, g); return;
          default: break;
        }
      }
      if constexpr (B16) {
        switch (v) {
          case 7: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 128, 208, SUB, 4, 1>), dim3(grid), dim3(256), 0, s, g); return;
          case 8: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 128, 208, BK2, 4, 1>), dim3(grid), dim3(256), 0, s, g); return;
          case 9: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 256, 208, SUB, 4, 1>), dim3(grid), dim3(256), 0, s, g); return;
          case 10: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 128, 128, BK2, 2, 1>), dim3(grid), dim3(256), 0, s, g); return;
          case 11: hipLaunchKernelGGL((gemm_nt_lds_kernel<TI, TO, 64, 208, BK2, 4, 1>), dim3(grid), dim3(256), 0, s, g); return;
          case 12: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 128, 208, 4, 3>), dim3(grid), dim3(256), 0, s, g); return;
          case 13: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 128, 208, 4, 4>), dim3(grid), dim3(256), 0, s, g); return;
          case 14: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 64, 208, 4, 4>), dim3(grid), dim3(256), 0, s, g); return;
          case 15: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 128, 128, 2, 4>), dim3(grid), dim3(256), 0, s, g); return;
          case 16: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 256, 128, 2, 3>), dim3(grid), dim3(256), 0, s, g); return;
          case 20: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 128, 208, 4, 3>), dim3(grid), dim3(256), 0, s, g); return;
          case 21: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 256, 128, 2, 3>), dim3(grid), dim3(256), 0, s, g); return;
          case 22: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 128, 128, 2, 4>), dim3(grid), dim3(256), 0, s, g); return;
          case 25: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 128, 224, 2, 3>), dim3(grid), dim3(256), 0, s, g); return;
          case 30: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 256, 208, 8, 3, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 31: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 256, 208, 8, 4, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 32: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 256, 224, 4, 3, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 33: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 256, 224, 4, 4, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 34: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 256, 208, 8, 2, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 35: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 256, 224, 4, 2, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 36: hipLaunchKernelGGL((gemm_nt_dma64_kernel<TI, TO, 128, 208, 8, 3, 8>), dim3(grid), dim3(512), 0, s, g); return;
          case 39: hipLaunchKernelGGL((gemm_rows_bf16_kernel<TO, 25, 4>), dim3(grid), dim3(512), 0, s, g); return;
          case 40: hipLaunchKernelGGL((gemm_rows_bf16_kernel<TO, 26, 4>), dim3(grid), dim3(512), 0, s, g); return;
          case 41: hipLaunchKernelGGL((gemm_rows_bf16_kernel<TO, 28, 4>), dim3(grid), dim3(512), 0, s, g); return;
          case 42:
            // column blocks in their own instance: the plain store path keeps the plain kernel's registers
            if (g.nblk > 0) hipLaunchKernelGGL((gemm_rankk_bf16_kernel<TO, 4, true>), dim3(grid), dim3(512), 0, s, g);
            else hipLaunchKernelGGL((gemm_rankk_bf16_kernel<TO, 4, false>), dim3(grid), dim3(512), 0, s, g);
            return;
          default:
            if (launch_dma_agpr(v, std::is_same<TO, double>::value, g, grid, s)) return;
            break;
        }
      }
      throw Error{RR_ERR_ARG, "LDS tile configuration not available for this precision"};
  }
}

#ifdef RR_VALIDATE
// Extent check of one GEMM launch (validation build): the highest scalar element each operand and
// output can touch, from the same index arithmetic as the kernels' loads and epilogues, against the
// allocation its pointer lies in. k is read in chunks of `kgran` scalars (a fragment of the register
// kernel, a 16-byte vector of the LDS-staged ones), so a row is read up to K rounded up to a chunk.
void validate_gemm(const GemmArgs& g, size_t si, size_t so, bool cplx, bool lds, const char* where) {
  if (g.M <= 0 || g.N <= 0) return;
  const long long cw = cplx ? 2 : 1;
  const long long kgran = lds ? (long long)(16 / si) : (si == 8 ? 4 : 8);
  const long long kblk = (long long)g.K * cw, kread = (kblk + kgran - 1) / kgran * kgran;
  const long long kb = std::max(1, g.kbatch);
  if (g.K > 0) {
    const long long a_end = ((long long)(g.nbatch - 1) * g.sAb + (kb - 1) * g.sAk + (long long)(g.M - 1) * g.lda) * cw + kread;
    const long long b_end = ((long long)(g.nbatch - 1) * g.sBb + (kb - 1) * g.sBk + (long long)(g.N - 1) * g.ldb) * cw + kread;
    rrv_check_range(g.A, (unsigned long long)a_end * si, "GEMM operand A", where);
    rrv_check_range(g.B, (unsigned long long)b_end * si, "GEMM operand B", where);
  }
  // Output: the offset of (m, n) is monotone inside each piece cut by mcut / ncut and by the row
  // and column blocks, so its maximum over a piece sits at the piece's last row and column.
  auto ends = [](long long lo, long long hi, long long blk, std::vector<long long>& out) {  // [lo, hi)
    if (hi <= lo) return;
    out.push_back(hi - 1);
    if (blk > 0)
      for (long long e = (lo / blk + 1) * blk - 1; e < hi - 1; e += blk) out.push_back(e);
  };
  std::vector<long long> ms, ns;
  if (g.mcut > 0 && g.mcut < g.M) {
    ends(0, g.mcut, g.mrow, ms);
    std::vector<long long> m2;
    ends(0, g.M - g.mcut, g.mrow, m2);
    for (long long v : m2) ms.push_back(v + g.mcut);
  } else {
    ends(0, g.M, g.mrow, ms);
  }
  if (g.ncut > 0 && g.ncut < g.N) {
    ends(0, g.ncut, g.nblk, ns);
    std::vector<long long> n2;
    ends(0, g.N - g.ncut, g.nblk, n2);
    for (long long v : n2) ns.push_back(v + g.ncut);
  } else {
    ends(0, g.N, g.nblk, ns);
  }
  const long long zoff = ((long long)(g.nbatch - 1) * g.sCb + (long long)(g.nsplit - 1) * g.sCz) * cw;
  long long c_end = -1, c2_end = -1;
  for (long long m : ms)
    for (long long n : ns) {
      long long mm = m, nn = n, off = 0;
      bool second = false;
      if (g.mcut > 0 && mm >= g.mcut) mm -= g.mcut, second = true;
      if (g.ncut > 0 && nn >= g.ncut) nn -= g.ncut, second = true;
      if (g.nblk > 0) {
        const long long b = nn / g.nblk;
        nn -= b * g.nblk;
        off += b * g.sCnb * cw;
      }
      if (g.mrow > 0) {
        const long long b = mm / g.mrow;
        mm -= b * g.mrow;
        off += b * g.sCmb;
      }
      off += (mm * g.ldcm + nn * g.ldcn) * cw + cw;  // one past the last scalar written
      long long& e = second ? c2_end : c_end;
      e = std::max(e, zoff + off);
    }
  if (c_end > 0) rrv_check_range(g.C, (unsigned long long)c_end * so, "GEMM output C", where);
  if (c2_end > 0) rrv_check_range(g.C2, (unsigned long long)c2_end * so, "GEMM second output C2", where);
  if (g.mrow > 0 && !lds) rrv_check_range(nullptr, 1, "GEMM row blocks on a kernel without them", where);
}
#endif

void run_gemm(Context& c, const GemmSpec& sp, GemmArgs g) {
  RR_REQUIRE(sp.cfg >= 0 && sp.cfg < kNumTiles, RR_ERR_ARG, "bad tile configuration");
  const bool use_lds = sp.lds >= 0 && !sp.cplx && g.kbatch == 0;
#ifdef RR_VALIDATE
  {
    char where[160];
    snprintf(where, sizeof(where), "run_gemm M=%d N=%d K=%d nbatch=%d nsplit=%d kbatch=%d prec=%d out=%d cplx=%d lds=%d",
             g.M, g.N, g.K, g.nbatch, g.nsplit, g.kbatch, sp.prec, sp.out, (int)sp.cplx, use_lds ? sp.lds : -1);
    validate_gemm(g, prec_size(sp.prec), prec_size(sp.out), sp.cplx, use_lds, where);
  }
#endif
  RR_REQUIRE(!use_lds || sp.lds < kNumLds, RR_ERR_ARG, "bad LDS tile configuration");
  const TileCfg t = kTiles[sp.cfg];
  const int BM = use_lds ? kLds[sp.lds].bm : 32 * t.wm, BN = use_lds ? kLds[sp.lds].bn : 32 * t.wn;
  if (use_lds && kLds[sp.lds].rows == 2) {
    RR_REQUIRE(sp.prec == P16 && sp.out == P64 && rankk_kernel_fits(g), RR_ERR_ARG,
               "persistent rank-k configuration does not fit this product");
    g.tiles_m = (g.M + BM - 1) / BM;
    g.tiles_n = (g.N + BN - 1) / BN;
    g.ntiles = g.tiles_m * g.tiles_n;
    if (g.kper <= 0) g.kper = g.K;
    c.last_ntiles = g.ntiles;
    g.zpage = c.zpage;
    const long long items = (long long)g.ntiles * g.nsplit;
    RR_REQUIRE(items < (1ll << 31), RR_ERR_ARG, "GEMM grid too large");
    const int grid = (int)std::max<long long>(1, std::min<long long>(items, c.num_cu));
    launch_lds<bf16_t, double>(sp.lds, g, grid, c.stream);
    RR_HIP(hipGetLastError());
    return;
  }
  if (use_lds && kLds[sp.lds].rows == 1) {
    // one tall product over the merged (batch, row) rows, persistent workgroups (one per CU)
    RR_REQUIRE(sp.prec == P16 && rows_kernel_fits(g, BN / 16, 4, prec_size(sp.out)), RR_ERR_ARG,
               "row-merged GEMM configuration does not fit this product");
    g.tiles_m = (int)(((long long)g.M * g.nbatch + BM - 1) / BM);
    g.tiles_n = (g.N + BN - 1) / BN;
    g.ntiles = g.tiles_m * g.tiles_n;
    if (g.kper <= 0) g.kper = g.K;
    c.last_ntiles = g.ntiles;
    g.zpage = c.zpage;
    const int grid = std::max(1, std::min(g.ntiles, c.num_cu));
    if (sp.out == P16) launch_lds<bf16_t, bf16_t>(sp.lds, g, grid, c.stream);
    else if (sp.out == P64) launch_lds<bf16_t, double>(sp.lds, g, grid, c.stream);
    else throw Error{RR_ERR_ARG, "unsupported GEMM precision combination"};
    RR_HIP(hipGetLastError());
    return;
  }
  g.tiles_m = (g.M + BM - 1) / BM;
  g.tiles_n = (g.N + BN - 1) / BN;
  if (g.tri) {
    RR_REQUIRE(BM == BN && g.M == g.N, RR_ERR_ARG, "triangular tiling needs square tiles and output");
    g.ntiles = g.tiles_m * (g.tiles_m + 1) / 2;
  } else {
    g.ntiles = g.tiles_m * g.tiles_n;
  }
  if (g.kper <= 0) g.kper = g.K * std::max(1, g.kbatch);
  const long long total = (long long)g.ntiles * g.nbatch * g.nsplit;
  if (total == 0 || g.M == 0 || g.N == 0) return;
  RR_REQUIRE(total < (1ll << 31), RR_ERR_ARG, "GEMM grid too large");
  const int grid = (int)total;
  c.last_ntiles = g.ntiles;
  g.zpage = c.zpage;
  if (use_lds) {
    if (sp.prec == P64 && sp.out == P64) launch_lds<double, double>(sp.lds, g, grid, c.stream);
    else if (sp.prec == P32 && sp.out == P32) launch_lds<float, float>(sp.lds, g, grid, c.stream);
    else if (sp.prec == P32 && sp.out == P64) launch_lds<float, double>(sp.lds, g, grid, c.stream);
    else if (sp.prec == P16 && sp.out == P16) launch_lds<bf16_t, bf16_t>(sp.lds, g, grid, c.stream);
    else if (sp.prec == P16 && sp.out == P64) launch_lds<bf16_t, double>(sp.lds, g, grid, c.stream);
    else throw Error{RR_ERR_ARG, "unsupported GEMM precision combination"};
    RR_HIP(hipGetLastError());
    return;
  }
  const bool cp = sp.cplx;
#define RR_DISPATCH(TI, TO)                                              \
  do {                                                                   \
    if (cp)                                                              \
      launch_gemm_cfg<TI, TO, true>(sp.cfg, g, grid, c.stream);          \
    else                                                                 \
      launch_gemm_cfg<TI, TO, false>(sp.cfg, g, grid, c.stream);         \
  } while (0)
  if (sp.prec == P64 && sp.out == P64) RR_DISPATCH(double, double);
  else if (sp.prec == P32 && sp.out == P32) RR_DISPATCH(float, float);
  else if (sp.prec == P32 && sp.out == P64) RR_DISPATCH(float, double);
  else if (sp.prec == P16 && sp.out == P16) RR_DISPATCH(bf16_t, bf16_t);
  else if (sp.prec == P16 && sp.out == P32) RR_DISPATCH(bf16_t, float);
  else if (sp.prec == P16 && sp.out == P64) RR_DISPATCH(bf16_t, double);
  else throw Error{RR_ERR_ARG, "unsupported GEMM precision combination"};
#undef RR_DISPATCH
  RR_HIP(hipGetLastError());
}

// Tile autotuner: for a large GEMM seen for the first time, every candidate configuration is run
// once on the real operands and timed with HIP events; the fastest is cached per shape. All
// candidates sum every output element in the same order, so the choice never changes a result.
void gemm_auto(Context& c, GemmSpec sp, const GemmArgs& g) {
  if (c.tile_override >= 0 || c.lds_override >= 0 || c.lds_override == -2) {
    if (c.tile_override >= 0) {
      const TileCfg& t = kTiles[c.tile_override];
      if (!g.tri || t.wm == t.wn) sp.cfg = c.tile_override;
      sp.lds = -1;
    }
    if (c.lds_override >= 0 && !sp.cplx && g.kbatch == 0) {
      const LdsCfg& v = kLds[c.lds_override];
      const bool fp64_wide = sp.prec == P64 && (c.lds_override == 2 || c.lds_override == 5 || c.lds_override == 6);
      const bool unavailable = (v.bf16_only && sp.prec != P16) || (v.fp64_only && sp.prec == P16) ||
                               (v.rows == 1 && (sp.out != P16 || !rows_kernel_fits(g, v.bn / 16, 4, prec_size(sp.out)))) ||
                               (v.rows == 2 && (sp.out != P64 || !rankk_kernel_fits(g)));
      if ((!g.tri || v.bm == v.bn) && !fp64_wide && !unavailable) sp.lds = c.lds_override;
    }
    if (c.lds_override == -2) sp.lds = -1;
    run_gemm(c, sp, g);
    return;
  }
  const double flops = 2.0 * g.M * g.N * (double)g.K * std::max(1, g.nbatch) * std::max(1, g.kbatch) * (sp.cplx ? 4 : 1);
  if (!c.autotune || flops < 2e10 || g.beta != 0.0) {
    run_gemm(c, sp, g);
    return;
  }
  std::vector<long long> key = {sp.prec, sp.out, sp.cplx, g.M, g.N, g.K, g.nbatch, g.nsplit, g.tri, g.kbatch, g.ncut, g.mcut, g.nblk};
  auto it = c.tuned.find(key);
  if (it != c.tuned.end()) {
    sp.cfg = it->second.first, sp.lds = it->second.second;
    run_gemm(c, sp, g);
    return;
  }
  if (c.arena) {  // allocation-free call: no timing events, the call site's default tile
    run_gemm(c, sp, g);
    return;
  }
  std::vector<std::pair<int, int>> cand;
  for (int cfg : {1, 2, 5}) {
    const TileCfg t = kTiles[cfg];
    if (g.tri && t.wm != t.wn) continue;
    cand.push_back({cfg, -1});
  }
  if (!sp.cplx && g.kbatch == 0)
    for (int v = 0; v < kNumLds; ++v) {
      if (g.tri && kLds[v].bm != kLds[v].bn) continue;
      if (sp.prec == P64 && (v == 2 || v == 5 || v == 6)) continue;  // these exceed the fp64 register budget
      if (kLds[v].bf16_only && sp.prec != P16) continue;
      if (kLds[v].fp64_only && sp.prec == P16) continue;
      if (kLds[v].rows == 1 && (sp.out != P16 || !rows_kernel_fits(g, kLds[v].bn / 16, 4, prec_size(sp.out)))) continue;
      if (kLds[v].rows == 2 && (sp.out != P64 || !rankk_kernel_fits(g))) continue;
      cand.push_back({1, v});
    }
  hipEvent_t e0, e1;
  RR_HIP(hipEventCreate(&e0));
  RR_HIP(hipEventCreate(&e1));
  auto time_one = [&](const std::pair<int, int>& cd) {
    GemmSpec t = sp;
    t.cfg = cd.first, t.lds = cd.second;
    RR_HIP(hipEventRecord(e0, c.stream));
    run_gemm(c, t, g);
    RR_HIP(hipEventRecord(e1, c.stream));
    RR_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    RR_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  };
  // Two timings of every candidate, then a third of those within 15% of the fastest; each candidate
  // is ranked by its best time (single timings of equal kernels differ by a few percent, and a first
  // launch can run beside other work still draining from the device).
  std::vector<float> tms(cand.size());
  float first = 1e30f;
  for (int rep = 0; rep < 2; ++rep)
    for (size_t i = 0; i < cand.size(); ++i) {
      const float t = time_one(cand[i]);
      tms[i] = rep == 0 ? t : std::min(tms[i], t);
      first = std::min(first, tms[i]);
    }
  float best = 1e30f;
  std::pair<int, int> pick = {sp.cfg, sp.lds};
  for (size_t i = 0; i < cand.size(); ++i) {
    if (tms[i] <= 1.15f * first) tms[i] = std::min(tms[i], time_one(cand[i]));
    if (tms[i] < best) best = tms[i], pick = cand[i];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.tuned[key] = pick;
  sp.cfg = pick.first, sp.lds = pick.second;
  run_gemm(c, sp, g);  // leaves the output of the chosen configuration (identical to every other)
}

size_t prec_size(int p) { return p == P64 ? 8 : (p == P32 ? 4 : 2); }

const void* tensor_data(const Tensor& t, int prec) {
  const void* p = prec == P64 ? (const void*)t.f64 : (prec == P32 ? (const void*)t.f32 : (const void*)t.b16);
  RR_REQUIRE(p != nullptr, RR_ERR_ARG, "tensor slot has no copy in the requested precision");
  return p;
}

// ================================================================================================
// Small FP64 products on device (recovery, embeddings): C = op(A) op(B), any op in {N, T, C}.
// Operands are packed into k-contiguous zero-padded buffers so the NT engine applies.
// ================================================================================================
// C[r][c] (row stride ldc) = alpha * sum_s part[s][r][c] + beta * C[r][c], part rows of width w.
__global__ void k_reduce_scale(const double* part, int nsplit, size_t stride, size_t rows, size_t w, size_t ldc,
                               double alpha, double beta, double* C) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < rows * w; i += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += part[z * stride + i];
    double* out = C + (i / w) * ldc + (i % w);
    *out = alpha * s + (beta != 0.0 ? beta * *out : 0.0);
  }
}

__global__ void k_pack_op(const double* in, size_t ld_in, char op, int cplx, size_t R, size_t K, size_t ld_out,
                          double* out) {
  // out row r (length ld_out) = op(in) row r, i.e. for op N: in[r][k]; for T/C: in[k][r].
  const size_t total = R * ld_out;
  const int cw = cplx ? 2 : 1;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t r = idx / ld_out, k = idx % ld_out;
    double re = 0.0, im = 0.0;
    if (k < K) {
      const size_t src = (op == 'N') ? r * ld_in + k : k * ld_in + r;
      re = in[cw * src];
      if (cplx) im = in[cw * src + 1];
      if (op == 'C') im = -im;
    }
    out[cw * idx] = re;
    if (cplx) out[cw * idx + 1] = im;
  }
}

// C (m x n, row-major, ldc) = op(A) (m x k) * op(B) (k x n); A/B device pointers with their own
// row-major leading dimensions. cplx: interleaved complex.
void dgemm_small(Context& c, bool cplx, char opa, char opb, size_t m, size_t n, size_t k, const double* A, size_t lda,
                 const double* B, size_t ldb, double* C, size_t ldc, double alpha, double beta) {
  if (m == 0 || n == 0) return;
  if (cplx && k > 0) {
    // Library zgemm on the column-major views (a row-major X is X^T column-major): C^T = op(B)^T op(A)^T,
    // where op(X)^T of the view is the view itself for 'N', its transpose for 'T' and its conjugate
    // transpose for 'C'. No packing or split-K partials (the packed path costs about twice as much
    // on the n <= 1000 products of the occupied-set recovery).
    auto rop = [](char o) {
      return o == 'N' ? rocblas_operation_none : (o == 'T' ? rocblas_operation_transpose : rocblas_operation_conjugate_transpose);
    };
    const rocblas_double_complex al = {alpha, 0.0}, be = {beta, 0.0};
    const rocblas_status st = rocblas_zgemm(
        c.blas, rop(opb), rop(opa), (rocblas_int)n, (rocblas_int)m, (rocblas_int)k, &al,
        reinterpret_cast<const rocblas_double_complex*>(B), (rocblas_int)ldb,
        reinterpret_cast<const rocblas_double_complex*>(A), (rocblas_int)lda, &be,
        reinterpret_cast<rocblas_double_complex*>(C), (rocblas_int)ldc);
    RR_REQUIRE(st == rocblas_status_success, RR_ERR_SOLVER, "rocblas zgemm failed");
    return;
  }
  const int cw = cplx ? 2 : 1;
  const size_t kp = roundup(std::max<size_t>(k, 1), 8);
  double* Ap = static_cast<double*>(wsbuf(c, WS_TMP8, m * kp * cw * sizeof(double)));
  double* Bp = static_cast<double*>(wsbuf(c, WS_TMP9, n * kp * cw * sizeof(double)));
  // A operand rows m of op(A); B operand rows n of op(B)^T: op(B)^T row j = column j of op(B).
  hipLaunchKernelGGL(k_pack_op, dim3(grid_for(m * kp)), dim3(256), 0, c.stream, A, lda, opa, (int)cplx, m, k, kp, Ap);
  // op(B)^T: if opb == 'N', row j of B^T = column j of B -> read in[k][j] => use op 'T'; 'T' -> read in[j][k] => 'N';
  // 'C' -> conj of in[j][k] (conj, no transpose) -> handled as 'N' + conj flag below.
  const char pb = (opb == 'N') ? 'T' : 'N';
  hipLaunchKernelGGL(k_pack_op, dim3(grid_for(n * kp)), dim3(256), 0, c.stream, B, ldb, pb, (int)cplx, n, k, kp, Bp);
  RR_HIP(hipGetLastError());
  GemmArgs g;
  g.A = Ap, g.lda = kp, g.B = Bp, g.ldb = kp, g.C = C, g.ldcm = ldc, g.ldcn = 1;
  g.M = (int)m, g.N = (int)n, g.K = (int)kp, g.alpha = alpha, g.beta = beta;
  g.conj = (cplx && opb == 'C') ? 2 : 0;
  GemmSpec sp;
  sp.prec = P64, sp.out = P64, sp.cplx = cplx, sp.cfg = (m >= 512 && n >= 512) ? 2 : (m >= 128 && n >= 128 ? 1 : 0);
  // Split K when the output alone gives too few workgroups (e.g. 200 x 200 x 1000 products).
  const int bm = 32 * kTiles[sp.cfg].wm;
  const size_t tiles = ((m + bm - 1) / bm) * ((n + bm - 1) / bm);
  const size_t want = std::max<size_t>(1, std::min<size_t>(64, 512 / std::max<size_t>(1, tiles)));
  const size_t ns = std::min<size_t>(want, std::max<size_t>(1, kp / 64));
  if (ns > 1) {
    const size_t kper = roundup((kp + ns - 1) / ns, 16);
    const int nsplit = (int)((kp + kper - 1) / kper);
    double* part = static_cast<double*>(wsbuf(c, WS_TMP9 + 50, (size_t)nsplit * m * n * cw * sizeof(double)));
    GemmArgs p = g;
    p.C = part, p.ldcm = n, p.sCz = m * n, p.kper = (int)kper, p.nsplit = nsplit, p.alpha = 1.0, p.beta = 0.0;
    run_gemm(c, sp, p);
    hipLaunchKernelGGL(k_reduce_scale, dim3(grid_for(m * n * cw)), dim3(256), 0, c.stream, part, nsplit, m * n * cw, m,
                       n * cw, ldc * cw, alpha, beta, C);
    RR_HIP(hipGetLastError());
    return;
  }
  run_gemm(c, sp, g);
}

// Real row-major D = alpha op(A) op(B) + beta C through the library GEMM on the column-major views
// (D^T = op(B)^T op(A)^T), one launch per product; C may be null (beta = 0) and D may differ from C.
// The library's atomics mode stays off (its default), so the summation order is fixed run to run.
// For the chains of small products (n_occ x n_occ) of the orbital alignment, where the packed
// kernels' extra launches cost more than the products.
void dgemm_lib(Context& c, char opa, char opb, size_t m, size_t n, size_t k, const double* A, size_t lda,
               const double* B, size_t ldb, double alpha, const double* C, size_t ldc, double beta, double* D,
               size_t ldd) {
  if (m == 0 || n == 0) return;
  auto rop = [](char o) { return o == 'N' ? rocblas_operation_none : rocblas_operation_transpose; };
  const rocblas_status st = rocblas_gemm_ex(
      c.blas, rop(opb), rop(opa), (rocblas_int)n, (rocblas_int)m, (rocblas_int)k, &alpha, B, rocblas_datatype_f64_r,
      (rocblas_int)ldb, A, rocblas_datatype_f64_r, (rocblas_int)lda, &beta, C ? C : D, rocblas_datatype_f64_r,
      (rocblas_int)(C ? ldc : ldd), D, rocblas_datatype_f64_r, (rocblas_int)ldd, rocblas_datatype_f64_r,
      rocblas_gemm_algo_standard, 0, 0);
  RR_REQUIRE(st == rocblas_status_success, RR_ERR_SOLVER, "rocblas gemm_ex failed");
}

// Symmetric / Hermitian eigensolve in place (row-major). On return, row j of `a` holds eigenvector
// j (components along the row) and w the eigenvalues in ascending order.
void eig_device(Context& c, bool cplx, size_t n, double* a, double* w) {
  if (n == 0) return;
  double* e = static_cast<double*>(wsbuf(c, WS_EIGE, n * sizeof(double)));
  rocblas_int* info = static_cast<rocblas_int*>(wsbuf(c, WS_INFO, sizeof(rocblas_int)));
  rocblas_status st;
  // Row-major symmetric == column-major; a Hermitian row-major array read column-major is conj(A),
  // whose eigenvectors are conj(v): row j of the column-major result is then conj(v_j) -> conjugate back.
  if (!cplx)
    st = rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, a, (rocblas_int)n, w, e, info);
  else
    st = rocsolver_zheevd(c.blas, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n,
                          reinterpret_cast<rocblas_double_complex*>(a), (rocblas_int)n, w, e, info);
  RR_REQUIRE(st == rocblas_status_success, RR_ERR_SOLVER, "rocsolver eigensolver failed");
  rocblas_int hinfo = 0;
  RR_HIP(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  RR_REQUIRE(hinfo == 0, RR_ERR_SOLVER, cplx ? "diagonalizeHermitian: eigensolver failed to converge"
                                              : "diagonalizeSymmetric: eigensolver failed to converge");
  if (cplx) {
    // conjugate imaginary parts
    hipLaunchKernelGGL(k_conj_inplace, dim3(grid_for(n * n)), dim3(256), 0, c.stream, a, n * n);
    RR_HIP(hipGetLastError());
  }
}


// ================================================================================================
// Phase timing (HIP events on the library stream)
// ================================================================================================
// Each timed phase records a pair of events around its launches; the pairs are resolved (waited for
// and summed) only when the totals are read (rr_timing) or too many are pending, so timing adds no
// host synchronization to the builds and does not serialise work on different streams.
struct PhaseTimer {
  struct Pending {
    int phase;
    hipEvent_t a, b;
  };
  Context& c;
  int phase;
  hipEvent_t a = nullptr, b = nullptr;
  // Shared by every handle: the switch is atomic, the precision of the build in progress is per
  // calling thread, and the totals and the pending event pairs are guarded by `mu`.
  static std::atomic<bool> enabled;
  static thread_local int cur_prec;  // precision of the build in progress (set by the entry points)
  static thread_local bool suppressed;  // no events in allocation-free calls
  static double ms[24];
  static long long cnt[24];
  static std::vector<Pending> pending;
  static std::mutex mu;
  PhaseTimer(Context& cc, int ph) : c(cc), phase(ph + 8 * cur_prec) {
    if (!enabled || suppressed) return;
    RR_HIP(hipEventCreate(&a));
    RR_HIP(hipEventCreate(&b));
    RR_HIP(hipEventRecord(a, c.stream));
  }
  ~PhaseTimer() {
    if (!enabled || !a) return;
    if (hipEventRecord(b, c.stream) == hipSuccess) {
      std::lock_guard<std::mutex> lock(mu);
      pending.push_back({phase, a, b});
      if (pending.size() > 4096) resolve_locked();
    } else {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  }
  static void resolve() {
    std::lock_guard<std::mutex> lock(mu);
    resolve_locked();
  }
  static void resolve_locked() {
    for (const Pending& p : pending) {
      float t = 0.f;
      if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
        ms[p.phase] += t;
        cnt[p.phase] += 1;
      }
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending.clear();
  }
};
std::atomic<bool> PhaseTimer::enabled{false};
thread_local int PhaseTimer::cur_prec = 0;
thread_local bool PhaseTimer::suppressed = false;
double PhaseTimer::ms[24] = {0};
long long PhaseTimer::cnt[24] = {0};
std::vector<PhaseTimer::Pending> PhaseTimer::pending;
std::mutex PhaseTimer::mu;
struct PrecScope {
  int saved;
  explicit PrecScope(int p) : saved(PhaseTimer::cur_prec) { PhaseTimer::cur_prec = (p >= 0 && p < 3) ? p : 0; }
  ~PrecScope() { PhaseTimer::cur_prec = saved; }
};
enum { PH_GAMMA = 0, PH_COUL = 1, PH_GEMM1 = 2, PH_GEMM2 = 3, PH_ASM = 4, PH_RECOVER = 5 };

// ================================================================================================
// Coulomb and exchange cores
// ================================================================================================
template <typename TB>
TB* tensor_ptr(const Tensor& t);
template <>
double* tensor_ptr<double>(const Tensor& t) { return t.f64; }
template <>
float* tensor_ptr<float>(const Tensor& t) { return t.f32; }
template <>
bf16_t* tensor_ptr<bf16_t>(const Tensor& t) { return t.b16; }

// gamma[P] (+)= <Bt_P, Dt> where Dt (n x ld, device, type of the tensor copy) is D^T padded.
template <typename TB>
void gamma_stream(Context& c, const Tensor& t, const TB* dDt, double* gamma, bool accumulate) {
  const TB* B = tensor_ptr<TB>(t);
  RR_REQUIRE(B, RR_ERR_STATE, "tensor slot has no copy in the requested precision");
  const size_t rowlen = t.n * t.ld;
  constexpr int R = 8;
  const size_t nrb = (t.naux + R - 1) / R;
  // Split the row so that the grid has >= ~2048 workgroups.
  size_t nsplit = std::max<size_t>(1, std::min<size_t>(64, (2048 + nrb - 1) / nrb));
  const size_t vn = 16 / sizeof(TB);
  size_t chunk = roundup((rowlen + nsplit - 1) / nsplit, 256 * vn);
  nsplit = (rowlen + chunk - 1) / chunk;
  double* gpart = static_cast<double*>(wsbuf(c, WS_GPART, nsplit * t.naux * sizeof(double)));
  {
    PhaseTimer pt(c, PH_GAMMA);
    hipLaunchKernelGGL((k_gamma<TB, R>), dim3((unsigned)nrb, (unsigned)nsplit), dim3(256), 0, c.stream, B, rowlen,
                       t.naux, dDt, chunk, gpart);
    RR_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_gamma_reduce, dim3(grid_for(t.naux)), dim3(256), 0, c.stream, gpart, (int)nsplit, t.naux, gamma,
                     accumulate ? 1 : 0);
  RR_HIP(hipGetLastError());
}

// jpart[s][e] partial sums of Jt = sum_P gamma_P Bt_P; returns the number of splits.
template <typename TB>
int coulomb_stream(Context& c, const Tensor& t, const double* gamma, double* jpart_out_or_null, double** jpart_used,
                   bool tri = false) {
  const TB* B = tensor_ptr<TB>(t);
  RR_REQUIRE(B, RR_ERR_STATE, "tensor slot has no copy in the requested precision");
  const size_t rowlen = t.n * t.ld;
  constexpr int U = 4;  // full rows (8 vectors per thread measured slower for FP64)
  constexpr int UT = CoulTri<TB>::U;
  const size_t vn = 16 / sizeof(TB);
  const size_t nvec = tri ? tri_vectors(t.n, t.ld, vn) : 0;
  const size_t per_wg = 256 * vn * U;
  const size_t nx = tri ? (nvec + 256 * UT - 1) / (256 * UT) : (rowlen + per_wg - 1) / per_wg;
  // split count: about 4096 workgroups for the triangle stream (measured best for bf16, flat for
  // fp64), 2048 for full rows
  const size_t target = tri ? 4096 : 2048;
  size_t nsplit = std::max<size_t>(1, std::min<size_t>(32, (target + nx - 1) / nx));
  nsplit = std::min(nsplit, std::max<size_t>(1, t.naux));
  const size_t pps = (t.naux + nsplit - 1) / nsplit;
  nsplit = (t.naux + pps - 1) / pps;
  double* jpart = jpart_out_or_null ? jpart_out_or_null
                                    : static_cast<double*>(wsbuf(c, WS_JPART, std::max<size_t>(1, nsplit) * rowlen * sizeof(double)));
  {
    PhaseTimer pt(c, PH_COUL);
    if (tri)
      hipLaunchKernelGGL((k_coulomb_tri<TB, UT, CoulTri<TB>::PU>), dim3((unsigned)std::max<size_t>(nx, 1), (unsigned)nsplit), dim3(256), 0,
                         c.stream, B, rowlen, t.ld, nvec, t.naux, gamma, pps, jpart);
    else
      hipLaunchKernelGGL((k_coulomb<TB, U>), dim3((unsigned)nx, (unsigned)nsplit), dim3(256), 0, c.stream, B, rowlen,
                         t.naux, gamma, pps, jpart);
    RR_HIP(hipGetLastError());
  }
  *jpart_used = jpart;
  return (int)nsplit;
}

int gemm1_cfg(Context& c, size_t ncols) {
  (void)c;
  if (ncols <= 32) return 0;   // 32 x 32
  if (ncols <= 64) return 1;   // 64 x 64
  if (ncols <= 128) return 4;  // 128 x 64 -> (4,2): BM=128, BN=64
  return 5;                    // 64 x 224
}
int gemm2_cfg(Context& c, size_t n) {
  (void)c;
  if (n <= 64) return 1;
  return 2;  // 128 x 128
}

// LDS-staged variants for the hot real GEMMs (-1: generic kernel). c.lds_override >= 0 forces one.
int gemm1_lds(Context& c, size_t m, size_t ncols) {
  (void)c;
  if (m < 64 || ncols < 64) return -1;
  return ncols <= 128 ? 0 : 1;
}
int gemm2_lds(Context& c, size_t n) {
  (void)c;
  if (n < 128) return -1;
  return 0;
}

// Number of k-splits of the rank-k exchange GEMM: a function of the problem shape only, so that
// the tile configuration never changes the summation order of any element.
int gemm2_splits(size_t n, size_t kdim, int slab) {
  const size_t t = (n + 127) / 128;
  const size_t tiles = std::max<size_t>(1, t * (t + 1) / 2);
  // at most 2048 workgroups (eight full rounds of one workgroup per CU): a count just above a
  // multiple of the CU count leaves a nearly empty last round of long-running tiles
  size_t ns = std::max<size_t>(1, 2048 / tiles);
  // at least 256 slabs of k per split: small products (the molecules of a batch) keep long k-loops
  ns = std::min<size_t>(ns, std::max<size_t>(1, kdim / (size_t)(slab * 256)));
  // a multiple of 16 from 16 up: the persistent rank-k kernel deals 16 (256 x 256) tiles of an
  // n <= 1024 product per split, so its work items fill whole rounds of the CUs
  if (ns >= 16) ns -= ns % 16;
  return (int)std::min<size_t>(ns, 256);
}

struct ExchangeOut {
  double* kpart = nullptr;
  int ksplit = 0;
  size_t kstride = 0;
  int ktile = 0;  // tri tile size (0: full G)
  int ksym = 0;   // full G: assemble G + G^T
};

// Real exchange in factor form on one tensor: W = Bt_P Lt^T (and W2 = Bt_P Rt^T), then
// K partials = W W^T (tri) or W W2^T (full). Lt/Rt: device fp64, r x ld_l.
// Coulomb contraction from the half-transformed factors: with W[nu][(P,i)] = (B_P^T L)[nu][i],
//   gamma_P = sum_(nu,i) W[nu][(P,i)] dfac[nu][i] (+ the same over W2 with dfac columns [rp, 2rp)).
// Pass 1: each thread owns 8 consecutive flat k = (P,i) positions (one aux row P, rp % 8 == 0) and
// sums over a slice of nu rows; pass 2 reduces the slices and the rp/8 groups of each P in order.
// Eight consecutive elements (16-byte aligned) widened to double. Plain loads: the half-transformed
// factors were just written and non-temporal loads measured slower here (fp64 1.1 -> 1.9 ms).
template <typename TW>
__device__ inline void load8(const TW* p, double* o) {
  constexpr int VN = Vec16<TW>::N;
#pragma unroll
  for (int v = 0; v < 8 / VN; ++v) Vec16<TW>::load(p + v * VN, o + v * VN);
}
template <typename TW>
__device__ __forceinline__ void k_gamma_w_partial_body(const TW* W, const TW* W2, size_t kdim, size_t rp, size_t n,
                                                         const double* dfac, size_t lddf, size_t rows, double* part, unsigned bx_, unsigned by_, unsigned gx_) {
  const size_t g8 = bx_ * (size_t)blockDim.x + threadIdx.x;  // group of 8 k
  const size_t k0 = g8 * 8;
  if (k0 >= kdim) return;
  const size_t i0 = k0 % rp;
  const size_t nu0 = by_ * rows, nu1 = std::min(n, nu0 + rows);
  double acc = 0.0;
  for (size_t nu = nu0; nu < nu1; ++nu) {
    const TW* w = W + nu * kdim + k0;
    const double* df = dfac + nu * lddf + i0;
    double wv[8];
    load8<TW>(w, wv);
#pragma unroll
    for (int q = 0; q < 8; ++q) acc += wv[q] * df[q];
    if (W2) {
      load8<TW>(W2 + nu * kdim + k0, wv);
#pragma unroll
      for (int q = 0; q < 8; ++q) acc += wv[q] * df[rp + q];
    }
  }
  part[by_ * (kdim / 8) + g8] = acc;
}
template <typename TW>
__global__ __launch_bounds__(256) void k_gamma_w_partial(const TW* W, const TW* W2, size_t kdim, size_t rp, size_t n,
                                                         const double* dfac, size_t lddf, size_t rows, double* part) {
  k_gamma_w_partial_body<TW>(W, W2, kdim, rp, n, dfac, lddf, rows, part, blockIdx.x, blockIdx.y, gridDim.x);
}

// The same partial sums as k_gamma_w_partial (identical values, same summation order per output),
// with each thread owning one group of 8 columns i for NPP consecutive aux rows P: the 8 weights of a
// row nu are loaded once for NPP rows of W, and every load is a 16-byte vector. Rows nu are taken in
// pairs through two register buffers: the weights of row nu are loaded first, then the W values of
// the next row, then row nu is summed, so the in-order load counter lets the next row's values stay
// in flight.
template <typename TW, int NPP, bool HAS_W2>
__device__ __forceinline__ void k_gamma_w_blocks_impl(const TW* W, const TW* W2, size_t kdim, size_t rp, size_t n,
                                                        size_t naux, const double* dfac, size_t lddf, size_t rows,
                                                        double* part, unsigned bx_, unsigned by_) {
  using V = Vec16<TW>;
  using raw = typename V::raw_t;
  constexpr int NV = 8 / V::N;  // 16-byte vectors per 8 values
  constexpr int NW = HAS_W2 ? 2 : 1;
  const size_t g8n = rp / 8, nblk = (naux + NPP - 1) / NPP;
  const size_t t = bx_ * (size_t)blockDim.x + threadIdx.x;
  if (t >= nblk * g8n) return;
  const size_t q = t % g8n, i0 = 8 * q, P0 = (t / g8n) * NPP;
  const size_t nu0 = by_ * rows, nu1 = std::min(n, nu0 + rows);
  size_t off[NPP];
#pragma unroll
  for (int p = 0; p < NPP; ++p) off[p] = std::min(P0 + p, naux - 1) * rp + i0;
  double acc[NPP];
#pragma unroll
  for (int p = 0; p < NPP; ++p) acc[p] = 0.0;
  auto fetch = [&](size_t nu, raw(&b)[NW][NPP][NV]) {
    const size_t r = nu * kdim;
#pragma unroll
    for (int p = 0; p < NPP; ++p)
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        b[0][p][v] = V::raw(W + r + off[p] + v * V::N);
        if constexpr (HAS_W2) b[NW - 1][p][v] = V::raw(W2 + r + off[p] + v * V::N);
      }
  };
  auto weights = [&](size_t nu, double* d, double* d2) {
    const double* df = dfac + nu * lddf + i0;
#pragma unroll
    for (int v = 0; v < 4; ++v) Vec16<double>::load(df + 2 * v, d + 2 * v);
    if constexpr (HAS_W2) {
#pragma unroll
      for (int v = 0; v < 4; ++v) Vec16<double>::load(df + rp + 2 * v, d2 + 2 * v);
    }
  };
  // (sums of aux rows past the last one run on its values and are not stored: no branch per lane)
  auto use = [&](const raw(&b)[NW][NPP][NV], const double* d, const double* d2) {
#pragma unroll
    for (int p = 0; p < NPP; ++p) {
      double w[8];
#pragma unroll
      for (int v = 0; v < NV; ++v) V::widen(b[0][p][v], w + v * V::N);
#pragma unroll
      for (int e = 0; e < 8; ++e) acc[p] += w[e] * d[e];
      if constexpr (HAS_W2) {
#pragma unroll
        for (int v = 0; v < NV; ++v) V::widen(b[NW - 1][p][v], w + v * V::N);
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[p] += w[e] * d2[e];
      }
    }
  };
  raw b0[NW][NPP][NV], b1[NW][NPP][NV];
  double d[8], d2[8];
  fetch(nu0, b0);
  size_t nu = nu0;
  for (; nu + 1 < nu1; nu += 2) {  // pairs of rows; the scheduling barriers keep the load order
    weights(nu, d, d2);
    fetch(nu + 1, b1);
    __builtin_amdgcn_sched_barrier(0);
    use(b0, d, d2);
    weights(nu + 1, d, d2);
    fetch(std::min(nu + 2, nu1 - 1), b0);
    __builtin_amdgcn_sched_barrier(0);
    use(b1, d, d2);
  }
  if (nu < nu1) {  // odd row count: the last row
    weights(nu, d, d2);
    use(b0, d, d2);
  }
#pragma unroll
  for (int p = 0; p < NPP; ++p)
    if (P0 + p < naux) part[by_ * (kdim / 8) + (P0 + p) * g8n + q] = acc[p];
}
template <typename TW, int NPP>
__device__ __forceinline__ void k_gamma_w_blocks_body(const TW* W, const TW* W2, size_t kdim, size_t rp, size_t n,
                                                        size_t naux, const double* dfac, size_t lddf, size_t rows,
                                                        double* part, unsigned bx_, unsigned by_, unsigned gx_) {
  if (W2) k_gamma_w_blocks_impl<TW, NPP, true>(W, W2, kdim, rp, n, naux, dfac, lddf, rows, part, bx_, by_);
  else k_gamma_w_blocks_impl<TW, NPP, false>(W, nullptr, kdim, rp, n, naux, dfac, lddf, rows, part, bx_, by_);
}
template <typename TW, int NPP>
__global__ __launch_bounds__(256) void k_gamma_w_blocks(const TW* W, const TW* W2, size_t kdim, size_t rp, size_t n,
                                                        size_t naux, const double* dfac, size_t lddf, size_t rows,
                                                        double* part) {
  k_gamma_w_blocks_body<TW, NPP>(W, W2, kdim, rp, n, naux, dfac, lddf, rows, part, blockIdx.x, blockIdx.y, gridDim.x);
}


// gamma[P] = sum over splits z and column groups q of part[z][P][q]: one wave per aux row P (lane l
// takes the pairs l, l + 64, ... of the z-major list, then a fixed butterfly), workgroup bx_ holding
// rows bx_ * waves .. (grid-stride over gx_ workgroups).
__device__ __forceinline__ void k_gamma_w_reduce_body(const double* part, int nsplit, size_t kdim, size_t rp, size_t naux, double* gamma, unsigned bx_, unsigned by_, unsigned gx_) {
  const unsigned gpp = (unsigned)(rp / 8), cnt = (unsigned)nsplit * gpp;
  const size_t ng = kdim / 8;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wpb = blockDim.x >> 6;
  for (size_t P = (size_t)bx_ * wpb + wave; P < naux; P += (size_t)gx_ * wpb) {
    double s = 0.0;
    for (unsigned idx = lane; idx < cnt; idx += 64) {
      const unsigned z = idx / gpp, q = idx - z * gpp;
      s += part[z * ng + P * gpp + q];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) gamma[P] = s;
  }
}
// workgroups of 256 (four aux rows each) for k_gamma_w_reduce
inline unsigned gamma_reduce_grid(size_t naux) { return (unsigned)std::max<size_t>(1, std::min<size_t>((naux + 3) / 4, 1u << 20)); }
__global__ void k_gamma_w_reduce(const double* part, int nsplit, size_t kdim, size_t rp, size_t naux, double* gamma) {
  k_gamma_w_reduce_body(part, nsplit, kdim, rp, naux, gamma, blockIdx.x, blockIdx.y, gridDim.x);
}

// Rows nu per partial sum of the gamma kernels (about 4096 workgroups per launch); a function of the
// shape and the W type only, so every path that forms gamma from W sums each value in the same order.
template <typename TW>
size_t gamma_rows(size_t kdim, size_t rp, size_t n, size_t naux) {
  size_t gx;
  if constexpr (std::is_same<TW, double>::value) gx = (kdim / 8 + 255) / 256;
  else gx = ((naux + 3) / 4 * (rp / 8) + 255) / 256;
  return std::max<size_t>(8, (n * gx + 4095) / 4096);
}

// gamma[P] = sum_(nu,i) W[nu][(P,i)] dfac[nu][i] (+ W2 with dfac columns [rp, 2rp)) on the stream.
template <typename TW>
void gamma_from_w(Context& c, const TW* W, const TW* W2, size_t kdim, size_t rp, size_t n, size_t naux,
                  const double* dfac, size_t lddf, double* gamma) {
  if constexpr (std::is_same<TW, double>::value) {
    // fp64 W: one 64-byte run per thread and row is already one full vector load per lane group;
    // the per-thread layout of k_gamma_w_partial streams it faster (measured)
    const size_t ng = kdim / 8, gx = (ng + 255) / 256;
    const size_t rows = gamma_rows<TW>(kdim, rp, n, naux);
    const int nsplit = (int)((n + rows - 1) / rows);
    double* part = static_cast<double*>(wsbuf(c, WS_GPART, (size_t)nsplit * ng * sizeof(double)));
    hipLaunchKernelGGL((k_gamma_w_partial<TW>), dim3((unsigned)gx, (unsigned)nsplit), dim3(256), 0, c.stream, W, W2, kdim,
                       rp, n, dfac, lddf, rows, part);
    hipLaunchKernelGGL(k_gamma_w_reduce, dim3(gamma_reduce_grid(naux)), dim3(256), 0, c.stream, part, nsplit, kdim, rp, naux, gamma);
    RR_HIP(hipGetLastError());
    return;
  }
  constexpr int NPP = 4;
  const size_t threads = ((naux + NPP - 1) / NPP) * (rp / 8);
  const size_t gx = (threads + 255) / 256;
  const size_t rows = gamma_rows<TW>(kdim, rp, n, naux);
  const int nsplit = (int)((n + rows - 1) / rows);
  double* part = static_cast<double*>(wsbuf(c, WS_GPART, (size_t)nsplit * (kdim / 8) * sizeof(double)));
  hipLaunchKernelGGL((k_gamma_w_blocks<TW, NPP>), dim3((unsigned)gx, (unsigned)nsplit), dim3(256), 0, c.stream, W, W2,
                     kdim, rp, n, naux, dfac, lddf, rows, part);
  hipLaunchKernelGGL(k_gamma_w_reduce, dim3(gamma_reduce_grid(naux)), dim3(256), 0, c.stream, part, nsplit, kdim, rp, naux, gamma);
  RR_HIP(hipGetLastError());
}

// gamma partials of nd <= ND densities that share one W (no W2), density k weighted by
// dfac[k][nu][i] (rows of length rp, densities sdf apart): the thread layout of k_gamma_w_blocks, with
// the NPP x 8 values of W a thread owns in row nu loaded once for every density. Each value is summed
// in the single-density kernels' order (rows nu of the slice, then the 8 columns), so the partials
// equal theirs bit for bit. part[k] starts spart elements after part[k - 1].
template <typename TW, int NPP, int ND>
__global__ __launch_bounds__(256) void k_gamma_w_multi(const TW* W, size_t kdim, size_t rp, size_t n, size_t naux,
                                                       const double* dfac, size_t sdf, int nd, size_t rows,
                                                       double* part, size_t spart) {
  const size_t g8n = rp / 8, nblk = (naux + NPP - 1) / NPP;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= nblk * g8n) return;
  const size_t q = t % g8n, i0 = 8 * q, P0 = (t / g8n) * NPP;
  const size_t nu0 = blockIdx.y * rows, nu1 = std::min(n, nu0 + rows);
  double acc[ND][NPP];
#pragma unroll
  for (int k = 0; k < ND; ++k)
#pragma unroll
    for (int p = 0; p < NPP; ++p) acc[k][p] = 0.0;
  for (size_t nu = nu0; nu < nu1; ++nu) {
    double w[NPP][8];
#pragma unroll
    for (int p = 0; p < NPP; ++p) {
      if (P0 + p < naux) {
        load8<TW>(W + nu * kdim + (P0 + p) * rp + i0, w[p]);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) w[p][e] = 0.0;
      }
    }
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      if (k < nd) {
        double d[8];
        const double* df = dfac + k * sdf + nu * rp + i0;
#pragma unroll
        for (int v = 0; v < 4; ++v) Vec16<double>::load(df + 2 * v, d + 2 * v);
#pragma unroll
        for (int p = 0; p < NPP; ++p)
#pragma unroll
          for (int e = 0; e < 8; ++e) acc[k][p] += w[p][e] * d[e];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < ND; ++k)
    if (k < nd) {
#pragma unroll
      for (int p = 0; p < NPP; ++p)
        if (P0 + p < naux) part[k * spart + blockIdx.y * (kdim / 8) + (P0 + p) * g8n + q] = acc[k][p];
    }
}

// gamma of density blockIdx.y from its partials (k_gamma_w_reduce on part + y spart, gamma + y naux)
__global__ void k_gamma_w_reduce_multi(const double* part, size_t spart, int nsplit, size_t kdim, size_t rp, size_t naux,
                                       double* gamma) {
  k_gamma_w_reduce_body(part + blockIdx.y * spart, nsplit, kdim, rp, naux, gamma + blockIdx.y * naux, blockIdx.x, 0,
                        gridDim.x);
}

// gamma_k[P] = sum_(nu,i) W[nu][(P,i)] dfac[k][nu][i] for nd <= 8 densities in one pass over W (the
// value gamma_from_w gives for each density alone, bit for bit).
template <typename TW>
void gamma_from_w_multi(Context& c, const TW* W, size_t kdim, size_t rp, size_t n, size_t naux, const double* dfac,
                        int nd, double* gamma) {
  constexpr int NPP = 4, ND = 8;
  RR_REQUIRE(nd >= 1 && nd <= ND, RR_ERR_INTERNAL, "gamma_from_w_multi: density count");
  const size_t threads = ((naux + NPP - 1) / NPP) * (rp / 8);
  const size_t gx = (threads + 255) / 256;
  const size_t rows = gamma_rows<TW>(kdim, rp, n, naux);
  const int nsplit = (int)((n + rows - 1) / rows);
  const size_t spart = (size_t)nsplit * (kdim / 8);
  double* part = static_cast<double*>(wsbuf(c, WS_GPART, (size_t)nd * spart * sizeof(double)));
  hipLaunchKernelGGL((k_gamma_w_multi<TW, NPP, ND>), dim3((unsigned)gx, (unsigned)nsplit), dim3(256), 0, c.stream, W,
                     kdim, rp, n, naux, dfac, n * rp, nd, rows, part, spart);
  hipLaunchKernelGGL(k_gamma_w_reduce_multi, dim3(gamma_reduce_grid(naux), (unsigned)nd), dim3(256), 0, c.stream, part,
                     spart, nsplit, kdim, rp, naux, gamma);
  RR_HIP(hipGetLastError());
}

// dfac (n x ncols) for the Coulomb contraction of a factor-form density: column block b of the
// half-transformed W is weighted by rows of fac[b] (r x ld_l, i.e. factor^T), scaled by scale[b].
__device__ __forceinline__ void k_build_dfac_body(const double* f0, const double* f1, size_t r, size_t ld_l, size_t n, size_t rp, int nblk,
                             double s0, double s1, double* out, unsigned bx_, unsigned by_, unsigned gx_) {
  const size_t cols = nblk * rp;
  for (size_t idx = bx_ * (size_t)blockDim.x + threadIdx.x; idx < n * cols; idx += (size_t)gx_ * blockDim.x) {
    const size_t nu = idx / cols, col = idx % cols, b = col / rp, i = col % rp;
    double v = 0.0;
    const double* f = b == 0 ? f0 : f1;
    if (i < r && f) v = (b == 0 ? s0 : s1) * f[i * ld_l + nu];
    out[idx] = v;
  }
}
__global__ void k_build_dfac(const double* f0, const double* f1, size_t r, size_t ld_l, size_t n, size_t rp, int nblk,
                             double s0, double s1, double* out) {
  k_build_dfac_body(f0, f1, r, ld_l, n, rp, nblk, s0, s1, out, blockIdx.x, blockIdx.y, gridDim.x);
}

template <typename TI>
ExchangeOut exchange_real(Context& c, const Tensor& t, int prec, const double* dLt, const double* dRt, size_t r,
                          size_t ld_l, const double* dfac = nullptr, double* gamma_out = nullptr,
                          hipEvent_t gamma_done = nullptr) {
  ExchangeOut out;
  const size_t n = t.n, ld = t.ld, naux = t.naux;
  const size_t rp = roundup(std::max<size_t>(r, 1), 8);
  const int nfac = dRt ? 2 : 1;
  // Pack the factor rows (zero-padded to rp rows and ld columns) in the operand type.
  TI* lt = static_cast<TI*>(wsbuf(c, WS_LT, nfac * rp * ld * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(rp * ld)), dim3(256), 0, c.stream, dLt, ld_l, 0, rp, r, n, ld, 1.0, lt);
  if (dRt)
    hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(rp * ld)), dim3(256), 0, c.stream, dRt, ld_l, 0, rp, r, n, ld, 1.0,
                       lt + rp * ld);
  RR_HIP(hipGetLastError());
  const size_t kdim = naux * rp;  // flat k of the rank-k update
  TI* W = static_cast<TI*>(wsbuf(c, WS_W, nfac * n * kdim * sizeof(TI)));
  TI* W2 = dRt ? W + n * kdim : nullptr;
  {
    PhaseTimer pt(c, PH_GEMM1);
    GemmArgs g;
    g.A = tensor_data(t, prec);
    g.lda = ld, g.sAb = n * ld;
    g.B = lt, g.ldb = ld, g.sBb = 0;
    g.C = W, g.ldcm = kdim, g.ldcn = 1, g.sCb = rp;
    g.M = (int)n, g.N = (int)(nfac * rp), g.K = (int)roundup(n, 8), g.nbatch = (int)naux;  // columns >= n are zero
    if (dRt) g.C2 = W2, g.ncut = (int)rp;
    GemmSpec sp;
    sp.prec = prec, sp.out = prec, sp.cfg = gemm1_cfg(c, nfac * rp);
    sp.lds = gemm1_lds(c, n, nfac * rp);
    gemm_auto(c, sp, g);
  }
  if (gamma_out) {
    PhaseTimer pt(c, PH_GAMMA);
    gamma_from_w<TI>(c, W, nfac == 2 ? W2 : nullptr, kdim, rp, n, naux, dfac, nfac * rp, gamma_out);
    if (gamma_done) RR_HIP(hipEventRecord(gamma_done, c.stream));
  }
  const int slab = prec == P64 ? 16 : 32;
  const int ns = gemm2_splits(n, kdim, slab);
  size_t kper = roundup((kdim + ns - 1) / ns, slab);
  const int nsplit = (int)((kdim + kper - 1) / kper);
  double* kpart = static_cast<double*>(wsbuf(c, WS_KPART, (size_t)nsplit * n * n * sizeof(double)));
  {
    PhaseTimer pt(c, PH_GEMM2);
    GemmArgs g;
    g.A = W, g.lda = kdim;
    g.B = dRt ? W2 : W, g.ldb = kdim;
    g.C = kpart, g.ldcm = n, g.ldcn = 1, g.sCz = n * n;
    g.M = (int)n, g.N = (int)n, g.K = (int)kdim, g.kper = (int)kper, g.nsplit = nsplit;
    g.tri = dRt ? 0 : 1;
    GemmSpec sp;
    sp.prec = prec, sp.out = P64, sp.cfg = gemm2_cfg(c, n);
    sp.lds = gemm2_lds(c, n);
    gemm_auto(c, sp, g);
    out.ktile = dRt ? 0 : 1;
  }
  out.kpart = kpart, out.ksplit = nsplit, out.kstride = n * n;
  return out;
}

// J partial sums for a real density D (device fp64, n x n): returns jpart and split count.
template <typename TB>
int coulomb_real(Context& c, const Tensor& t, const double* dD, double** jpart, bool tri) {
  TB* dt = static_cast<TB*>(wsbuf(c, WS_DT, t.n * t.ld * sizeof(TB)));
  hipLaunchKernelGGL((k_pack<TB>), dim3(grid_for(t.n * t.ld)), dim3(256), 0, c.stream, dD, t.n, 1, t.n, t.n, t.n, t.ld, 1.0,
                     dt);
  RR_HIP(hipGetLastError());
  double* gamma = static_cast<double*>(wsbuf(c, WS_GAMMA, std::max<size_t>(1, t.naux) * sizeof(double)));
  gamma_stream<TB>(c, t, dt, gamma, false);
  return coulomb_stream<TB>(c, t, gamma, nullptr, jpart, tri);
}

// tri: write only the upper triangle of Jt (symmetric tensors; read back with RealAsm::jtri).
int coulomb_real_prec(Context& c, const Tensor& t, int prec, const double* dD, double** jpart, bool tri = false) {
  if (prec == P64) return coulomb_real<double>(c, t, dD, jpart, tri);
  if (prec == P32) return coulomb_real<float>(c, t, dD, jpart, tri);
  return coulomb_real<bf16_t>(c, t, dD, jpart, tri);
}

ExchangeOut exchange_real_prec(Context& c, const Tensor& t, int prec, const double* dLt, const double* dRt, size_t r,
                               size_t ld_l, const double* dfac = nullptr, double* gamma_out = nullptr,
                               hipEvent_t gamma_done = nullptr) {
  if (prec == P64) return exchange_real<double>(c, t, prec, dLt, dRt, r, ld_l, dfac, gamma_out, gamma_done);
  if (prec == P32) return exchange_real<float>(c, t, prec, dLt, dRt, r, ld_l, dfac, gamma_out, gamma_done);
  return exchange_real<bf16_t>(c, t, prec, dLt, dRt, r, ld_l, dfac, gamma_out, gamma_done);
}


// Coulomb stream only (gamma given): returns split partials of Jt.
int coulomb_from_gamma(Context& c, const Tensor& t, int prec, const double* gamma, double** jpart, bool tri = false) {
  if (prec == P64) return coulomb_stream<double>(c, t, gamma, nullptr, jpart, tri);
  if (prec == P32) return coulomb_stream<float>(c, t, gamma, nullptr, jpart, tri);
  return coulomb_stream<bf16_t>(c, t, gamma, nullptr, jpart, tri);
}

// The exchange with gamma taken from the half-transform, and the Coulomb stream of that gamma run
// on the auxiliary stream, concurrently with the rank-k update (the stream reads B, the update reads
// W), for fp64 / fp32 operands. Falls back to one stream when the context has no auxiliary stream.
ExchangeOut exchange_and_coulomb(Context& c, const Tensor& t, int prec, const double* dLt, const double* dRt, size_t r,
                                 size_t ld_l, const double* dfac, double* gamma, bool want_j, bool tri, double** jpart,
                                 int* jsplit) {
  // bf16: the rank-k update is bound by L2 traffic that the concurrent stream slows down as much as
  // it hides (measured); fp64 / fp32 updates are matrix-core bound and hide the stream.
  const bool overlap = want_j && c.aux && prec != P16;
  ExchangeOut ex = exchange_real_prec(c, t, prec, dLt, dRt, r, ld_l, dfac, gamma, overlap ? c.ev_fork : nullptr);
  *jsplit = 0;
  if (!want_j) return ex;
  if (!overlap) {
    *jsplit = coulomb_from_gamma(c, t, prec, gamma, jpart, tri);
    return ex;
  }
  RR_HIP(hipStreamWaitEvent(c.aux, c.ev_fork, 0));
  hipStream_t main = c.stream;
  c.stream = c.aux;
  try {
    *jsplit = coulomb_from_gamma(c, t, prec, gamma, jpart, tri);
  } catch (...) {
    c.stream = main;
    throw;
  }
  c.stream = main;
  RR_HIP(hipEventRecord(c.ev_join, c.aux));
  RR_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
  return ex;
}

void real_assemble(Context& c, const Tensor& t, double* jpart, int jsplit, const ExchangeOut* ex, double kfac,
                   const double* dh, double* dF, double* dJ, double* dK, bool jtri) {
  PhaseTimer pt(c, PH_ASM);
  RealAsm a;
  a.n = t.n, a.ld = t.ld, a.jtri = jtri ? 1 : 0;
  // Reduce the split partials first with coalesced reads, so that the assembly's transposed reads
  // touch one matrix instead of every partial.
  if (jpart && jsplit > 1) {
    double* jr = static_cast<double*>(wsbuf(c, WS_TMP8 + 400, t.n * t.ld * sizeof(double)));
    hipLaunchKernelGGL(k_reduce_parts, dim3(grid_for(t.n * t.ld)), dim3(256), 0, c.stream, jpart, jsplit, t.n * t.ld,
                       t.n * t.ld, jr);
    jpart = jr, jsplit = 1;
  }
  a.jpart = jpart, a.jsplit = jsplit, a.jstride = t.n * t.ld;
  if (ex && ex->kpart) {
    a.kpart = ex->kpart, a.ksplit = ex->ksplit, a.kstride = ex->kstride, a.ktile = ex->ktile, a.ksym = ex->ksym;
    if (ex->ksplit > 1) {
      double* kr = static_cast<double*>(wsbuf(c, WS_TMP8 + 401, t.n * t.n * sizeof(double)));
      hipLaunchKernelGGL(k_reduce_parts, dim3(grid_for(t.n * t.n)), dim3(256), 0, c.stream, ex->kpart, ex->ksplit,
                         ex->kstride, t.n * t.n, kr);
      a.kpart = kr, a.ksplit = 1;
    }
  }
  a.kfac = kfac, a.h = dh, a.F = dF, a.J = dJ, a.K = dK;
  hipLaunchKernelGGL(k_real_asm, dim3(grid_for(t.n * t.n)), dim3(256), 0, c.stream, a);
  RR_HIP(hipGetLastError());
}

Tensor& get_slot(Context& c, int slot) {
  RR_REQUIRE(slot >= 0 && slot < kSlots, RR_ERR_ARG, "tensor slot out of range");
  Tensor& t = c.slot[slot];
  RR_REQUIRE(t.f64 || t.f32 || t.b16, RR_ERR_STATE, "tensor slot is empty");
  return t;
}

void check_prec(int prec) { RR_REQUIRE(prec == P64 || prec == P32 || prec == P16, RR_ERR_ARG, "bad precision"); }

// Output helper: device pointer if the caller's buffer is device memory, else a workspace buffer
// that is copied back by finish_out().
OutBuf out_buf(Context& c, int id, double* user, size_t count) {
  OutBuf o;
  o.user = user, o.count = count;
  if (!user) return o;
  if (is_device_ptr(user)) {
#ifdef RR_VALIDATE
    rrv_check_range(user, count * sizeof(double), "caller output array (declared extent)", "out_buf");
#endif
    o.dev = user;
  } else {
    o.dev = static_cast<double*>(wsbuf(c, id, std::max<size_t>(1, count) * sizeof(double)));
    o.copy = true;
  }
  return o;
}
void finish_out(Context& c, const OutBuf& o) {
  if (o.copy && o.count) RR_HIP(hipMemcpyAsync(o.user, o.dev, o.count * sizeof(double), hipMemcpyDeviceToHost, c.stream));
}
const double* in_buf(Context& c, int id, const double* user, size_t count) {
  if (!user) return nullptr;
  if (is_device_ptr(user)) {
#ifdef RR_VALIDATE
    rrv_check_range(user, count * sizeof(double), "caller input array (declared extent)", "in_buf");
#endif
    return user;
  }
  return to_device<double>(c, id, user, count);
}


// Occupied-orbital recovery (NonRelFockMatrixRi.cpp:51-64, X2C_FockMatrixRi.cpp:65-95,
// UkbFockMatrixRi.cpp:192-209): eigendecompose Peff = X^H (S P S) X * scale (FP64, rocSOLVER),
// keep eigenvalues > 0.5 (the top of the ascending spectrum). Returns n_occ; on return
// `rows` points at n_occ eigenvector rows (length nx, complex when cplx) and `occ_host` holds
// their eigenvalues.
// ---- spectral-projector recovery ---------------------------------------------------------------
// out = (M - shift I) * scale over an n x n array (interleaved complex when cw = 2).
__global__ void k_shift_scale(const double* m, size_t n, int cw, double shift, double scale, double* out) {
  const size_t total = n * n * cw;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t e = i / cw;
    double v = m[i];
    if (i % cw == 0 && e / n == e % n) v -= shift;
    out[i] = v * scale;
  }
}

// Per-block partial sums of |T - I|^2 (sub_identity) or |T|^2 over an n x n array.
__global__ void k_fro_partial(const double* t, size_t n, int cw, int sub_identity, double* part) {
  __shared__ double sh[256];
  const size_t total = n * n * cw;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t e = i / cw;
    double v = t[i];
    if (sub_identity && i % cw == 0 && e / n == e % n) v -= 1.0;
    acc += v * v;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// *trace = Re tr (I + X) / 2 (one workgroup).
__global__ void k_half_trace(const double* x, size_t n, int cw, double* trace) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (size_t d = threadIdx.x; d < n; d += blockDim.x) acc += 0.5 * (x[(d * n + d) * cw] + 1.0);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *trace = sh[0];
}

// X <- (I + X) / 2 in place.
__global__ void k_projector(double* x, size_t n, int cw) {
  const size_t total = n * n * cw;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t e = i / cw;
    x[i] = 0.5 * (x[i] + ((i % cw == 0 && e / n == e % n) ? 1.0 : 0.0));
  }
}

// Deterministic test matrix: uniform values in [-1, 1) from a splitmix64 hash of (seed, index).
__global__ void k_fill_uniform(double* out, size_t count, unsigned long long seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

// out (cols x rows) = in^T (plain transpose) of a rows x cols array.
__global__ void k_transpose_any(const double* in, size_t rows, size_t cols, int cw, double* out) {
  const size_t total = rows * cols;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / cols, col = i % cols;
    for (int p = 0; p < cw; ++p) out[(col * rows + r) * cw + p] = in[i * cw + p];
  }
}

// Zero the strictly lower triangle of a row-major n x n array.
__global__ void k_zero_strict_lower(double* a, size_t n, int cw) {
  const size_t total = n * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
    if (i % n < i / n)
      for (int p = 0; p < cw; ++p) a[i * cw + p] = 0.0;
}

namespace {
constexpr int kFroBlocks = 256;

double fro2(Context& c, const double* t, size_t n, int cw, bool sub_identity, const char* tag) {
  double* part = static_cast<double*>(wsbuf(c, WS_T4 + 800, kFroBlocks * sizeof(double)));
  hipLaunchKernelGGL(k_fro_partial, dim3(kFroBlocks), dim3(256), 0, c.stream, t, n, cw, (int)sub_identity, part);
  RR_HIP(hipGetLastError());
  double h[kFroBlocks];
  RR_HIP(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (double v : h) s += v;
  RR_CONTROL(tag, s);
  (void)tag;
  return s;
}

// Cholesky of a row-major Hermitian r x r array in place: its upper triangle becomes R with G = R^H R
// (the column-major lower factor of G^T = conj(G), read row-major).
void chol_upper(Context& c, bool cplx, size_t r, double* g, rocblas_int* info) {
  rocblas_status st = cplx ? rocsolver_zpotrf(c.blas, rocblas_fill_lower, (rocblas_int)r,
                                              reinterpret_cast<rocblas_double_complex*>(g), (rocblas_int)r, info)
                           : rocsolver_dpotrf(c.blas, rocblas_fill_lower, (rocblas_int)r, g, (rocblas_int)r, info);
  RR_REQUIRE(st == rocblas_status_success, RR_ERR_SOLVER, "rocsolver Cholesky failed");
}

// Y (rows x r, row-major) <- Y R^{-1} for the upper factor R of chol_upper (column-major: R^T Y^T = Y^T).
void solve_right_upper(Context& c, bool cplx, size_t rows, size_t r, const double* g, double* y) {
  rocblas_status st;
  if (cplx) {
    const rocblas_double_complex one = {1.0, 0.0};
    st = rocblas_ztrsm(c.blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                       (rocblas_int)r, (rocblas_int)rows, &one, reinterpret_cast<const rocblas_double_complex*>(g),
                       (rocblas_int)r, reinterpret_cast<rocblas_double_complex*>(y), (rocblas_int)r);
  } else {
    const double one = 1.0;
    st = rocblas_dtrsm(c.blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                       (rocblas_int)r, (rocblas_int)rows, &one, g, (rocblas_int)r, y, (rocblas_int)r);
  }
  RR_REQUIRE(st == rocblas_status_success, RR_ERR_SOLVER, "rocblas triangular solve failed");
}
}  // namespace

// Factor of the occupied part of a Hermitian M (nx x nx) without an eigensolver:
//   Pi = step(M - 1/2) from the Newton-Schulz sign iteration X <- X (3 - X^2) / 2 started at
//   (M - I/2) / s, s >= ||M - I/2||_2 (s = ||(M - I/2)^8||_F^(1/8)); r = tr Pi;
//   Q = orthonormal basis of range(Pi Omega) for a fixed pseudo-random nx x r Omega (Cholesky QR, then
//   Newton-Schulz orthogonalization steps);
//   weighted: F = Q R_H^H with Q^H M Q = R_H^H R_H, so F F^H = Pi M Pi = sum_{lambda > 1/2} lambda v v^H;
//   otherwise F = Q, F F^H = Pi = sum_{lambda > 1/2} v v^H.
// Any F with the same F F^H gives the same exchange, so F replaces the reference's sqrt(occ)-scaled
// eigenvectors. Writes F^T (r x nx) to rows_out and returns r, or returns -1 (no result) when the
// iteration does not settle in a fixed number of steps (an eigenvalue at or near 1/2), when tr Pi is
// not an integer or a Cholesky factorization fails; the caller then uses the eigensolver.
__global__ void k_axpby(size_t count, double a, const double* x, double b, const double* y, double* out);

long projector_factor(Context& c, bool cplx, size_t nx, const double* M, bool weighted, double* rows_out) {
  const int cw = cplx ? 2 : 1;
  const size_t nn = nx * nx * cw;
  double* X = static_cast<double*>(wsbuf(c, WS_T4 + 801, nn * sizeof(double)));
  double* T = static_cast<double*>(wsbuf(c, WS_T4 + 802, nn * sizeof(double)));
  // Scale of A = M - I/2 from its Frobenius powers: hi = ||A^8||_F^(1/8) >= ||A||_2 (at most
  // nx^(1/16) above it), lo = (||A^8||_F / ||A^4||_F)^(1/4) <= ||A||_2 (Cauchy-Schwarz; equal to it
  // when every eigenvalue of A has the same magnitude, as for an idempotent density). The start
  // (M - I/2) / s with s = max(lo, 1.01 hi / sqrt 3) keeps every eigenvalue below sqrt 3 in
  // magnitude, where the iteration converges to the right sign, and is exactly +-1 for an
  // idempotent density.
  hipLaunchKernelGGL(k_shift_scale, dim3(grid_for(nn)), dim3(256), 0, c.stream, M, nx, cw, 0.5, 1.0, X);
  RR_HIP(hipGetLastError());
  dgemm_small(c, cplx, 'N', 'N', nx, nx, nx, X, nx, X, nx, T, nx);  // A^2
  dgemm_small(c, cplx, 'N', 'N', nx, nx, nx, T, nx, T, nx, X, nx);  // A^4
  const double a4 = std::sqrt(fro2(c, X, nx, cw, false, "projector_scale"));
  dgemm_small(c, cplx, 'N', 'N', nx, nx, nx, X, nx, X, nx, T, nx);  // A^8
  const double a8 = std::sqrt(fro2(c, T, nx, cw, false, "projector_scale"));
  if (!(a8 > 0.0) || !std::isfinite(a8) || !(a4 > 0.0) || !std::isfinite(a4)) return -1;
  const double hi = std::pow(a8, 0.125), lo = std::pow(a8 / a4, 0.25);
  const double s = std::max(lo, 1.01 * hi / std::sqrt(3.0));
  hipLaunchKernelGGL(k_shift_scale, dim3(grid_for(nn)), dim3(256), 0, c.stream, M, nx, cw, 0.5, 1.0 / s, X);
  RR_HIP(hipGetLastError());
  constexpr int kMaxSteps = 40;
  bool done = false;
  for (int it = 0; it < kMaxSteps && !done; ++it) {
    dgemm_small(c, cplx, 'N', 'N', nx, nx, nx, X, nx, X, nx, T, nx);  // X^2
    const double e = std::sqrt(fro2(c, T, nx, cw, true, "projector_residual"));  // ||X^2 - I||_F
    if (!std::isfinite(e) || e > 4.0 * nx) return -1;
    done = e < 1e-9;  // one more step leaves an error at rounding level
    dgemm_small(c, cplx, 'N', 'N', nx, nx, nx, X, nx, T, nx, X, nx, -0.5, 1.5);  // X (3 - X^2) / 2
  }
  if (!done) return -1;
  double* tr = static_cast<double*>(wsbuf(c, WS_T4 + 803, sizeof(double)));
  hipLaunchKernelGGL(k_half_trace, dim3(1), dim3(256), 0, c.stream, X, nx, cw, tr);
  hipLaunchKernelGGL(k_projector, dim3(grid_for(nn)), dim3(256), 0, c.stream, X, nx, cw);
  RR_HIP(hipGetLastError());
  double trace = 0.0;
  RR_HIP(hipMemcpyAsync(&trace, tr, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  RR_CONTROL("projector_trace", trace);
  const double rr_ = std::nearbyint(trace);
  if (!(std::fabs(trace - rr_) < 1e-6) || rr_ < 0.0 || rr_ > (double)nx) return -1;
  const size_t r = (size_t)rr_;
  if (r == 0) return 0;
  // basis of range(Pi): Y = Pi Omega, one Cholesky-QR pass and Newton-Schulz steps
  double* om = static_cast<double*>(wsbuf(c, WS_T4 + 804, nx * r * cw * sizeof(double)));
  double* Y = static_cast<double*>(wsbuf(c, WS_T4 + 805, nx * r * cw * sizeof(double)));
  double* G = static_cast<double*>(wsbuf(c, WS_T4 + 806, r * r * cw * sizeof(double)));
  rocblas_int* info = static_cast<rocblas_int*>(wsbuf(c, WS_T4 + 807, 4 * sizeof(rocblas_int)));
  RR_HIP(hipMemsetAsync(info, 0, 4 * sizeof(rocblas_int), c.stream));
  hipLaunchKernelGGL(k_fill_uniform, dim3(grid_for(nx * r * cw)), dim3(256), 0, c.stream, om, nx * r * cw,
                     0x5EEDF00DULL + nx);
  RR_HIP(hipGetLastError());
  dgemm_small(c, cplx, 'N', 'N', nx, r, nx, X, nx, om, r, Y, r);
  const char ct = cplx ? 'C' : 'T';
  dgemm_small(c, cplx, ct, 'N', r, r, nx, Y, r, Y, r, G, r);  // Y^H Y
  chol_upper(c, cplx, r, G, info);
  solve_right_upper(c, cplx, nx, r, G, Y);
  // Cholesky QR leaves ||Q^H Q - I|| at about cond(Y)^2 eps; Newton-Schulz steps Q <- Q (3 I - Q^H Q) / 2
  // square that error (in place of a second Cholesky QR pass: two small products per step instead of
  // a product, a factorization and a triangular solve).
  for (int it = 0;; ++it) {
    dgemm_small(c, cplx, ct, 'N', r, r, nx, Y, r, Y, r, G, r);  // S = Q^H Q
    const double e = std::sqrt(fro2(c, G, r, cw, true, "projector_residual"));  // ||S - I||_F
    if (!(e < 0.1) || it == 4) return -1;  // also a failed factorization (NaN)
    dgemm_small(c, cplx, 'N', 'N', nx, r, r, Y, r, G, r, om, r);  // Q S
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(nx * r * cw)), dim3(256), 0, c.stream, nx * r * cw, 1.5, Y, -0.5, om, Y);
    RR_HIP(hipGetLastError());
    if (e < 1e-9) break;  // the step just taken leaves about e^2 < 1e-18: rounding level
  }
  double* F = Y;
  if (weighted) {
    dgemm_small(c, cplx, 'N', 'N', nx, r, nx, M, nx, Y, r, om, r);  // M Q
    dgemm_small(c, cplx, ct, 'N', r, r, nx, Y, r, om, r, G, r);     // H = Q^H M Q
    chol_upper(c, cplx, r, G, info + 2);
    hipLaunchKernelGGL(k_zero_strict_lower, dim3(grid_for(r * r)), dim3(256), 0, c.stream, G, r, cw);
    RR_HIP(hipGetLastError());
    dgemm_small(c, cplx, 'N', ct, nx, r, r, Y, r, G, r, om, r);  // F = Q R_H^H
    F = om;
  }
  rocblas_int hinfo[4] = {0, 0, 0, 0};
  RR_HIP(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  if (hinfo[0] || hinfo[1] || hinfo[2]) return -1;
  hipLaunchKernelGGL(k_transpose_any, dim3(grid_for(nx * r)), dim3(256), 0, c.stream, F, nx, r, cw, rows_out);
  RR_HIP(hipGetLastError());
  return (long)r;
}

// Occupied set of a density (X2C_FockMatrixRi.cpp:65-95, NonRelFockMatrixRi.cpp:51-69,
// UkbFockMatrixRi.cpp:185-211): M = scale X^H (S P S) X; the occupied vectors are the eigenvectors of M
// with eigenvalue above 1/2. On return *rows holds n_occ rows of length nx (row i = occupied vector i)
// and occ_host their weights: eigenvalues of M (eigensolver), or ones for the rows of a factor F with
// F F^H = sum lambda v v^H (projector path, `weighted`), or ones with F F^H = sum v v^H (unweighted:
// callers that use only the vectors, such as the NON_REL build with its fixed sqrt 2).
size_t recover_occupied(Context& c, bool cplx, size_t n, size_t nx, const double* dS, const double* dP, const double* dX,
                        double scale, const double** rows, std::vector<double>& occ_host, bool weighted) {
  PhaseTimer pt(c, PH_RECOVER);
  const int cw = cplx ? 2 : 1;
  const char T = cplx ? 'C' : 'T';
  double* t1 = static_cast<double*>(wsbuf(c, WS_T1, n * n * cw * sizeof(double)));
  double* t2 = static_cast<double*>(wsbuf(c, WS_T2, n * n * cw * sizeof(double)));
  double* t3 = static_cast<double*>(wsbuf(c, WS_T3, n * std::max<size_t>(nx, 1) * cw * sizeof(double)));
  double* t4 = static_cast<double*>(wsbuf(c, WS_T4, std::max<size_t>(nx * nx, 1) * cw * sizeof(double)));
  dgemm_small(c, cplx, 'N', 'N', n, n, n, dS, n, dP, n, t1, n);                 // S P
  dgemm_small(c, cplx, 'N', 'N', n, n, n, t1, n, dS, n, t2, n, scale);          // S P S * scale
  dgemm_small(c, cplx, 'N', 'N', n, nx, n, t2, n, dX, nx, t3, nx);              // (SPS) X
  dgemm_small(c, cplx, T, 'N', nx, nx, n, dX, nx, t3, nx, t4, nx);              // X^H (SPS) X
  if (nx == 0) return 0;
  if (c.recovery_method == 0) {
    const long r = projector_factor(c, cplx, nx, t4, weighted, t3);
    if (r >= 0) {
      ++c.n_projector;
      occ_host.assign((size_t)r, 1.0);
      *rows = t3;
      return (size_t)r;
    }
  }
  double* w = static_cast<double*>(wsbuf(c, WS_EIGW, std::max<size_t>(nx, 1) * sizeof(double)));
  eig_device(c, cplx, nx, t4, w);
  std::vector<double> wh(nx);
  RR_HIP(hipMemcpyAsync(wh.data(), w, nx * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  ++c.n_eigensolver;
  size_t first = nx;
  while (first > 0 && wh[first - 1] > 0.5) --first;
  occ_host.assign(wh.begin() + first, wh.end());
  *rows = t4 + first * nx * cw;
  return nx - first;
}

__global__ void k_scale_rows_sqrt(double* a, size_t rows, size_t cols, int cw, const double* occ) {
  const size_t total = rows * cols * cw;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
    a[i] *= sqrt(occ[i / (cols * cw)]);
}


// c = a*x + b*y elementwise
__global__ void k_axpby(size_t count, double a, const double* x, double b, const double* y, double* out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    out[i] = a * x[i] + b * y[i];
}
// out = 1.5 X - 0.5 X (X^T X) helper pieces: Y = 1.5*X - 0.5*Z
// max over i of |a_i - delta_i| for the m x m elements, four independent loads in flight per thread
__device__ inline double max_dev_identity(const double* a, size_t m, double scale) {
  const size_t mm = m * m, step = blockDim.x;
  double d[4] = {0.0, 0.0, 0.0, 0.0};
  size_t i = threadIdx.x;
  for (; i + 3 * step < mm; i += 4 * step) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t j = i + u * step;
      d[u] = fmax(d[u], fabs(a[j] * scale - ((j / m) == (j % m) ? 1.0 : 0.0)));
    }
  }
  for (; i < mm; i += step) d[0] = fmax(d[0], fabs(a[i] * scale - ((i / m) == (i % m) ? 1.0 : 0.0)));
  return fmax(fmax(d[0], d[1]), fmax(d[2], d[3]));
}
__device__ inline double block_max(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, sh[w]);
  return r;
}
// For the symmetric m x m matrix a = X^T X (one workgroup, 1024 threads): out[0] = b = max_i sum_j |a_ij|
// (summed along columns, which the symmetry makes the row sums, so the loads run along rows),
// out[1] = max_ij |a_ij / b - delta_ij|.
__global__ __launch_bounds__(1024) void k_align_scale(const double* a, size_t m, double* out) {
  __shared__ double sh[16];
  __shared__ double part[1024];
  double e = 0.0;
  if (m <= blockDim.x) {
    // G thread groups per column, group g summing the rows i = g mod G; the G partial sums of a
    // column are then added in group order
    const size_t G = blockDim.x / m, g = threadIdx.x / m, j = threadIdx.x % m;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    if (g < G) {
      size_t i = g;
      for (; i + 3 * G < m; i += 4 * G) {
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] += fabs(a[(i + u * G) * m + j]);
      }
      for (; i < m; i += G) r[0] += fabs(a[i * m + j]);
    }
    part[threadIdx.x] = (r[0] + r[1]) + (r[2] + r[3]);
    __syncthreads();
    if (threadIdx.x < m) {
      double cs = 0.0;
      for (size_t h = 0; h < G; ++h) cs += part[h * m + threadIdx.x];
      e = cs;
    }
  } else {
    for (size_t j = threadIdx.x; j < m; j += blockDim.x) {
      double r[4] = {0.0, 0.0, 0.0, 0.0};
      size_t i = 0;
      for (; i + 3 < m; i += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] += fabs(a[(i + u) * m + j]);
      }
      for (; i < m; ++i) r[0] += fabs(a[i * m + j]);
      e = fmax(e, (r[0] + r[1]) + (r[2] + r[3]));
    }
  }
  const double b = block_max(e, sh);
  const double d = block_max(max_dev_identity(a, m, 1.0 / b), sh);
  if (threadIdx.x == 0) out[0] = b, out[1] = d;
}
// ||X^T X - I||_max into *err (one workgroup, 1024 threads)
__global__ __launch_bounds__(1024) void k_orth_err(const double* xtx, size_t m, double* err) {
  __shared__ double sh[16];
  const double e = block_max(max_dev_identity(xtx, m, 1.0), sh);
  if (threadIdx.x == 0) *err = e;
}

// Incremental build for D = 2 C C^T between two orbital sets (rr_ri_jk_delta):
//   Q = polar(C_new^T C_old) (Newton-Schulz, FP64), C' = C_new Q (same density as C_new, aligned to
//   C_old), dC = C' - C_old, M = C_old + dC/2, so D_new - D_old = 2 (M dC^T + dC M^T) exactly;
//   dJ = J(dD); dK = G + G^T with G = sum_P (B_P^T sqrt2 M)(B_P^T sqrt2 dC)^T.
void delta_build(Context& c, const Tensor& t, int prec, size_t n, size_t no, const double* dA, const double* dB,
                 double* dJ, double* dK) {
  if (no == 0) {
    if (dJ) hipLaunchKernelGGL(k_zero, dim3(grid_for(n * n)), dim3(256), 0, c.stream, dJ, n * n);
    if (dK) hipLaunchKernelGGL(k_zero, dim3(grid_for(n * n)), dim3(256), 0, c.stream, dK, n * n);
    return;
  }
  double* q = static_cast<double*>(wsbuf(c, WS_T1, no * no * sizeof(double)));
  double* z = static_cast<double*>(wsbuf(c, WS_T2, no * no * sizeof(double)));
  double* y = static_cast<double*>(wsbuf(c, WS_T3, no * no * sizeof(double)));
  double* err = static_cast<double*>(wsbuf(c, WS_TMP8 + 100, 2 * sizeof(double)));
  // C_new^T C_old (k = n: the split-k product fills the device better than one library launch)
  dgemm_small(c, false, 'T', 'N', no, no, n, dA, no, dB, no, q, no);
  // The polar factor is invariant to a positive scale; dividing by an upper bound of the largest
  // singular value puts every singular value in (0, 1], where the Newton-Schulz step X (3 - X^T X) / 2
  // converges to it (the orbitals need not be orthonormal in the plain metric, e.g. C^T S C = I, so
  // the singular values of C_new^T C_old are not bounded by 1 themselves). The bound is
  // sigma_max^2 <= ||X^T X||_inf (max row sum of |X^T X|): exact for an orthogonal X, whose alignment
  // then needs no step at all (the Frobenius norm, sqrt(n_occ) times too large there, cost about ten).
  dgemm_lib(c, 'T', 'N', no, no, no, q, no, q, no, 1.0, nullptr, 0, 0.0, z, no);  // X^T X
  // bound and the orthogonality error of the scaled X in one pass and one read-back
  hipLaunchKernelGGL(k_align_scale, dim3(1), dim3(1024), 0, c.stream, z, no, err);
  double be[2] = {0.0, 0.0};
  RR_HIP(hipMemcpyAsync(be, err, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  RR_CONTROL("align_scale", be[0]);
  const double bnd = be[0];
  const double qn = std::sqrt(bnd);
  bool aligned = qn > 0.0 && std::isfinite(qn);
  if (aligned) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(no * no)), dim3(256), 0, c.stream, no * no, 1.0 / qn, q, 0.0, q, q);
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(no * no)), dim3(256), 0, c.stream, no * no, 1.0 / bnd, z, 0.0, z, z);
    aligned = false;
    for (int it = 0; it < 120 && !aligned; ++it) {
      if (it > 0) dgemm_lib(c, 'T', 'N', no, no, no, q, no, q, no, 1.0, nullptr, 0, 0.0, z, no);  // X^T X
      if (it == 0) {
        RR_CONTROL("align_residual", be[1]);
        const double e = be[1];
        if (!std::isfinite(e)) break;
        aligned = e < 1e-13;  // an orthogonal overlap: one step below leaves rounding level
      } else if (it % 3 == 0) {
        hipLaunchKernelGGL(k_orth_err, dim3(1), dim3(1024), 0, c.stream, z, no, err);
        double e = 0;
        RR_HIP(hipMemcpyAsync(&e, err, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        RR_HIP(hipStreamSynchronize(c.stream));
        RR_CONTROL("align_residual", e);
        if (!std::isfinite(e)) break;
        aligned = e < 1e-13;  // the step below then leaves rounding level
      }
      dgemm_lib(c, 'N', 'N', no, no, no, q, no, z, no, -0.5, q, no, 1.5, y, no);  // 1.5 X - 0.5 X (X^T X)
      std::swap(q, y);
    }
  }
  double* cp = static_cast<double*>(wsbuf(c, WS_T4, n * no * sizeof(double)));
  // C' = C_new Q; without an alignment (occupied spaces with a nearly orthogonal direction) C' = C_new:
  // the difference identity below holds for any C', the alignment only keeps dC small.
  if (aligned)
    dgemm_lib(c, 'N', 'N', n, no, no, dA, no, q, no, 1.0, nullptr, 0, 0.0, cp, no);
  else
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n * no)), dim3(256), 0, c.stream, n * no, 1.0, dA, 0.0, dA, cp);
  double* dc = static_cast<double*>(wsbuf(c, WS_COCC, 2 * n * no * sizeof(double)));
  double* m = dc + n * no;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n * no)), dim3(256), 0, c.stream, n * no, 1.0, cp, -1.0, dB, dc);   // dC
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n * no)), dim3(256), 0, c.stream, n * no, 1.0, dB, 0.5, dc, m);     // M
  RR_HIP(hipGetLastError());
  // Factors sqrt2*M and sqrt2*dC as rows (Lt, Rt). The Coulomb contraction is fused into the
  // half-transform: gamma_P(dD) = sum_(nu,i) W_M[nu][(P,i)] sqrt2 dC[nu][i] + W_dC[nu][(P,i)] sqrt2 M[nu][i].
  double* lt = static_cast<double*>(wsbuf(c, WS_T1 + 200, 2 * no * n * sizeof(double)));
  hipLaunchKernelGGL((k_pack<double>), dim3(grid_for(no * n)), dim3(256), 0, c.stream, m, no, 1, no, no, n, n,
                     std::sqrt(2.0), lt);
  hipLaunchKernelGGL((k_pack<double>), dim3(grid_for(no * n)), dim3(256), 0, c.stream, dc, no, 1, no, no, n, n,
                     std::sqrt(2.0), lt + no * n);
  RR_HIP(hipGetLastError());
  const size_t rp = roundup(no, 8);
  double* dfac = nullptr;
  double* gamma = nullptr;
  if (dJ) {
    dfac = static_cast<double*>(wsbuf(c, WS_DENS2, n * 2 * rp * sizeof(double)));
    hipLaunchKernelGGL(k_build_dfac, dim3(grid_for(n * 2 * rp)), dim3(256), 0, c.stream, lt + no * n, lt, no, n, n, rp, 2,
                       1.0, 1.0, dfac);
    RR_HIP(hipGetLastError());
    gamma = static_cast<double*>(wsbuf(c, WS_GAMMA, std::max<size_t>(1, t.naux) * sizeof(double)));
  }
  double* jpart = nullptr;
  int jsplit = 0;
  ExchangeOut ex = exchange_and_coulomb(c, t, prec, lt, lt + no * n, no, n, dfac, gamma, dJ != nullptr, t.symmetric,
                                        &jpart, &jsplit);
  ex.ksym = 1;
  real_assemble(c, t, jpart, jsplit, dK ? &ex : nullptr, 0.0, nullptr, nullptr, dJ, dK, t.symmetric);
}


// ================================================================================================
// Many densities sharing the tensor and the left factor, D_k = L R_k^T (the trial densities of the
// orbital-Hessian sigma products: L the occupied orbitals, R_k = C_virt T_k^T). W_L = B L is built
// once; gamma_k = <W_L, R_k> for 8 densities per pass over W_L; the R_k are half-transformed G at a
// time in one launch (column blocks of the output) and K_k = sum_P W_L,P W_Rk,P^T for a stack of
// them per rank-k launch (column blocks again); J_k = sum_P gamma_k[P] B_P for all densities in
// passes of 8 over the tensor.
// ================================================================================================
enum { WS_MU_LT = 200000, WS_MU_WL, WS_MU_RT, WS_MU_WR, WS_MU_DF, WS_MU_GP, WS_MU_GAM, WS_MU_KP, WS_MU_JP,
       WS_MU_IN_L, WS_MU_IN_R, WS_MU_OUT_J, WS_MU_OUT_K };

// gt[P][k] = gamma[k][P] for k < nk, 0 for nk <= k < NK: the NK weights of one aux row side by side.
template <int NK>
__global__ void k_gamma_rows(const double* gamma, size_t gstride, int nk, size_t naux, double* gt) {
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < naux * NK; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t P = idx / NK;
    const int k = (int)(idx % NK);
    gt[idx] = k < nk ? gamma[k * gstride + P] : 0.0;
  }
}

// jpart[split][k][e] = sum over the split's aux rows of gt[P][k] Bt[P][e] for NK densities (gt from
// k_gamma_rows; columns past the densities in use are zero and their sums are not stored), over the
// upper-triangle vectors (tri) or every vector of the row. Each thread owns one 16-byte vector. The
// weights of up to CH aux rows at a time are staged in the LDS (read as broadcasts); the tensor rows
// go through two register buffers of PU rows: the next PU rows are loaded (row index clamped to the
// chunk, so every load is unconditional and in bounds) while the current PU rows are used, about
// eight fp64 multiply-adds per loaded value. The sums run over P in order.
template <typename TB, int NK, int PU>
__global__ __launch_bounds__(256) void k_coulomb_multi(const TB* B, size_t rowlen, size_t ld, size_t nvec, int tri,
                                                       size_t naux, const double* gt, int nk, size_t pps,
                                                       double* jpart) {
  using V = Vec16<TB>;
  using raw = typename V::raw_t;
  constexpr int VN = V::N, CH = 256;
  __shared__ double gsh[CH * NK];
  const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
  const bool live = v < nvec;  // threads past the last vector load a valid one and store nothing
  const size_t vv = live ? v : nvec - 1;
  const size_t split = blockIdx.y;
  const size_t pb = split * pps, pe = std::min(naux, pb + pps);
  const size_t e = tri ? tri_offset(vv, ld / VN, VN, ld) : vv * VN;
  const TB* col = B + e;
  double acc[NK][VN];
#pragma unroll
  for (int k = 0; k < NK; ++k)
#pragma unroll
    for (int q = 0; q < VN; ++q) acc[k][q] = 0.0;
  for (size_t c0 = pb; c0 < pe; c0 += CH) {
    const size_t c1 = std::min(pe, c0 + CH);
    __syncthreads();  // the previous chunk's weights are no longer read
    for (size_t i = threadIdx.x; i < (c1 - c0) * NK; i += 256) gsh[i] = gt[c0 * NK + i];
    __syncthreads();
    auto fetch = [&](size_t P, raw(&buf)[PU]) {
#pragma unroll
      for (int h = 0; h < PU; ++h) buf[h] = V::raw_nt(col + std::min(P + h, c1 - 1) * rowlen);
    };
    auto use = [&](size_t P, const raw(&buf)[PU]) {
#pragma unroll
      for (int h = 0; h < PU; ++h) {
        if (P + h < c1) {
          const double* g = gsh + (P + h - c0) * NK;
          double w[VN];
          V::widen(buf[h], w);
#pragma unroll
          for (int k = 0; k < NK; ++k) {
            const double gk = g[k];
#pragma unroll
            for (int q = 0; q < VN; ++q) acc[k][q] += gk * w[q];
          }
        }
      }
    };
    raw b0[PU], b1[PU];
    fetch(c0, b0);
    for (size_t P = c0; P < c1; P += 2 * PU) {
      fetch(P + PU, b1);
      use(P, b0);
      fetch(P + 2 * PU, b0);
      use(P + PU, b1);
    }
  }
  if (!live) return;
  double* out = jpart + split * (size_t)NK * rowlen;
#pragma unroll
  for (int k = 0; k < NK; ++k)
    if (k < nk) {
#pragma unroll
      for (int q = 0; q < VN; ++q) out[k * rowlen + e + q] = acc[k][q];
    }
}

// J_k[mu][nu] = sum_s jpart[s][k][.] at (min, max) (tri) or (nu, mu) of the transposed layout, in
// 32 x 32 tiles of the stored layout (workgroup x: tile, y: density): the split sums are read along
// stored rows and written through the LDS both as they lie and transposed, so every global access
// runs along a row. Each value sums the splits in order.
__global__ __launch_bounds__(256) void k_multi_jasm(const double* jpart, int nsplit, size_t sstride, size_t rowlen,
                                                    size_t ld, size_t n, int tri, double* J) {
  __shared__ double T[32][33];
  const size_t nb = (n + 31) / 32;
  const size_t br = blockIdx.x / nb, bc = blockIdx.x % nb;  // stored rows br*32.., columns bc*32..
  if (tri && br > bc) return;  // uniform per workgroup, before any barrier
  const double* src = jpart + blockIdx.y * rowlen;
  double* Jk = J + blockIdx.y * n * n;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const bool diag = tri && br == bc;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int a = ty + 8 * j;
    const size_t r = br * 32 + a, c = bc * 32 + tx;
    double s = 0.0;
    if (r < n && c < n && (!diag || a <= tx))
      for (int z = 0; z < nsplit; ++z) s += src[z * sstride + r * ld + c];
    T[a][tx] = s;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int a = ty + 8 * j;
    if (tri) {  // as stored: J[r][c] (the lower half of a diagonal tile from its upper half)
      const size_t r = br * 32 + a, c = bc * 32 + tx;
      if (r < n && c < n) Jk[r * n + c] = (diag && a > tx) ? T[tx][a] : T[a][tx];
    }
    if (!diag) {  // transposed: J[c][r] = stored (r, c)
      const size_t row = bc * 32 + a, c = br * 32 + tx;
      if (row < n && c < n) Jk[row * n + c] = T[tx][a];
    }
  }
}

template <typename TI>
void multi_jk(Context& c, const Tensor& t, int prec, const double* dL, const double* dR, size_t r, size_t nd,
              double* dJ, double* dK) {
  const size_t n = t.n, ld = t.ld, naux = t.naux;
  const size_t rp = roundup(std::max<size_t>(r, 1), 8), kdim = naux * rp;
  const size_t G = std::min(nd, std::max<size_t>(1, 416 / rp));  // densities per half-transform launch
  const int tri = t.symmetric ? 1 : 0;
  // W_L[nu][(P,i)] = (B_P^T L)[nu][i]
  TI* lt = static_cast<TI*>(wsbuf(c, WS_MU_LT, rp * ld * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(rp * ld)), dim3(256), 0, c.stream, dL, r, 1, rp, r, n, ld, 1.0, lt);
  RR_HIP(hipGetLastError());
  TI* WL = static_cast<TI*>(wsbuf(c, WS_MU_WL, n * kdim * sizeof(TI)));
  auto half = [&](const TI* rows, size_t ncols, TI* out, int nblk, long long sCnb) {
    PhaseTimer pt(c, PH_GEMM1);
    GemmArgs g;
    g.A = tensor_data(t, prec), g.lda = ld, g.sAb = n * ld;
    g.B = rows, g.ldb = ld;
    g.C = out, g.ldcm = kdim, g.ldcn = 1, g.sCb = rp;
    g.M = (int)n, g.N = (int)ncols, g.K = (int)roundup(n, 8), g.nbatch = (int)naux;
    g.nblk = nblk, g.sCnb = sCnb;
    GemmSpec sp;
    sp.prec = prec, sp.out = prec, sp.cfg = gemm1_cfg(c, ncols);
    sp.lds = gemm1_lds(c, n, ncols);
    gemm_auto(c, sp, g);
  };
  half(lt, rp, WL, 0, 0);
  double* gam = dJ ? static_cast<double*>(wsbuf(c, WS_MU_GAM, nd * naux * sizeof(double))) : nullptr;
  if (dJ) {
    // gamma_k = <W_L, R_k> for up to 8 densities per pass over W_L
    constexpr size_t GD = 8;
    double* dfac = static_cast<double*>(wsbuf(c, WS_MU_DF, GD * n * rp * sizeof(double)));
    for (size_t k0 = 0; k0 < nd; k0 += GD) {
      const size_t nk = std::min(GD, nd - k0);
      PhaseTimer pt(c, PH_GAMMA);
      for (size_t j = 0; j < nk; ++j)
        hipLaunchKernelGGL((k_pack<double>), dim3(grid_for(n * rp)), dim3(256), 0, c.stream, dR + (k0 + j) * n * r, r, 0, n,
                           n, r, rp, 1.0, dfac + j * n * rp);
      RR_HIP(hipGetLastError());
      gamma_from_w_multi<TI>(c, WL, kdim, rp, n, naux, dfac, (int)nk, gam + k0 * naux);
    }
  }
  if (dK) {
    const int slab = prec == P64 ? 16 : 32;
    const int ns = gemm2_splits(n, kdim, slab);
    const size_t kper = roundup((kdim + ns - 1) / ns, slab);
    const int nsplit = (int)((kdim + kper - 1) / kper);
    // K_k = sum_P W_L,P W_Rk,P^T for GR densities per launch: their W_R blocks lie stacked in WR
    // (G per half-transform launch), and the product writes each K_k as a column block. Same
    // splits as the single-density build, so every K_k is summed in its order. Up to 16 densities
    // while the stacked factors stay within 8 GiB and the partial outputs within 4 GiB: the W_L
    // panels are read once for all of them, and an orbital count that is not a multiple of the
    // tile edge wastes tile area in one dimension only.
    const size_t wr_one = n * kdim * sizeof(TI), kp_one = (size_t)nsplit * n * n * sizeof(double);
    size_t GR = G;
    while (GR + G <= std::min<size_t>(nd, 16) && (GR + G) * wr_one <= (8ull << 30) && (GR + G) * kp_one <= (4ull << 30))
      GR += G;
    TI* rt = static_cast<TI*>(wsbuf(c, WS_MU_RT, G * rp * ld * sizeof(TI)));
    TI* WR = static_cast<TI*>(wsbuf(c, WS_MU_WR, GR * wr_one));
    double* kpart = static_cast<double*>(wsbuf(c, WS_MU_KP, GR * kp_one));
    for (size_t g0 = 0; g0 < nd; g0 += GR) {
      const size_t gr = std::min(GR, nd - g0);
      for (size_t h0 = 0; h0 < gr; h0 += G) {
        const size_t gs = std::min(G, gr - h0);
        for (size_t j = 0; j < gs; ++j)
          hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(rp * ld)), dim3(256), 0, c.stream, dR + (g0 + h0 + j) * n * r, r, 1,
                             rp, r, n, ld, 1.0, rt + j * rp * ld);
        RR_HIP(hipGetLastError());
        half(rt, gs * rp, WR + h0 * n * kdim, (int)rp, (long long)(n * kdim));
      }
      {
        PhaseTimer pt(c, PH_GEMM2);
        GemmArgs g;
        g.A = WL, g.lda = kdim;
        g.B = WR, g.ldb = kdim;
        g.C = kpart, g.ldcm = n, g.ldcn = 1, g.sCz = gr * n * n;
        g.nblk = (int)n, g.sCnb = (long long)(n * n);
        g.M = (int)n, g.N = (int)(gr * n), g.K = (int)kdim, g.kper = (int)kper, g.nsplit = nsplit;
        GemmSpec sp;
        sp.prec = prec, sp.out = P64, sp.cfg = gemm2_cfg(c, n);
        sp.lds = gemm2_lds(c, n);
        gemm_auto(c, sp, g);
      }
      hipLaunchKernelGGL(k_reduce_parts, dim3(grid_for(gr * n * n)), dim3(256), 0, c.stream, kpart, nsplit, gr * n * n,
                         gr * n * n, dK + g0 * n * n);
      RR_HIP(hipGetLastError());
    }
  }
  if (dJ) {
    constexpr int NK = 8, PU = 4;
    const TI* B = static_cast<const TI*>(tensor_data(t, prec));
    const size_t vn = 16 / sizeof(TI), rowlen = n * ld;
    const size_t nvec = tri ? tri_vectors(n, ld, vn) : rowlen / vn;
    const size_t nx = std::max<size_t>(1, (nvec + 255) / 256);
    size_t js = std::max<size_t>(1, std::min<size_t>(32, (2048 + nx - 1) / nx));
    js = std::min(js, std::max<size_t>(1, naux));
    const size_t pps = (naux + js - 1) / js;
    js = (naux + pps - 1) / pps;
    double* jp = static_cast<double*>(wsbuf(c, WS_MU_JP, js * NK * rowlen * sizeof(double)));
    double* gt = static_cast<double*>(wsbuf(c, WS_MU_GP, naux * NK * sizeof(double)));
    const size_t nb = (n + 31) / 32;
    for (size_t k0 = 0; k0 < nd; k0 += NK) {
      const int nk = (int)std::min<size_t>(NK, nd - k0);
      {
        PhaseTimer pt(c, PH_COUL);
        hipLaunchKernelGGL((k_gamma_rows<NK>), dim3(grid_for(naux * NK)), dim3(256), 0, c.stream, gam + k0 * naux, naux, nk,
                           naux, gt);
        hipLaunchKernelGGL((k_coulomb_multi<TI, NK, PU>), dim3((unsigned)nx, (unsigned)js), dim3(256), 0, c.stream, B,
                           rowlen, ld, nvec, tri, naux, gt, nk, pps, jp);
        RR_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(k_multi_jasm, dim3((unsigned)(nb * nb), (unsigned)nk), dim3(256), 0, c.stream, jp, (int)js,
                         NK * rowlen, rowlen, ld, n, tri, dJ + k0 * n * n);
      RR_HIP(hipGetLastError());
    }
  }
}

void multi_jk_prec(Context& c, const Tensor& t, int prec, const double* dL, const double* dR, size_t r, size_t nd,
                   double* dJ, double* dK) {
  if (prec == P64) multi_jk<double>(c, t, prec, dL, dR, r, nd, dJ, dK);
  else if (prec == P32) multi_jk<float>(c, t, prec, dL, dR, r, nd, dJ, dK);
  else multi_jk<bf16_t>(c, t, prec, dL, dR, r, nd, dJ, dK);
}

void drop_plan(GraphPlan& p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  p = GraphPlan();
}


// ================================================================================================
// Complex (multi-flavor) exchange: X2C spin blocks and the four UKB flavors
// (X2C_FockMatrixRi.cpp:97-146, UkbFockMatrixRi.cpp:84-160)
// ================================================================================================
// GEMM1 operand rows for two flavors sharing one tensor: row (fl*2rp + 2i + c), column a, holds
//   c_t_f(i,a) = conj(C(off_f + a, i)) sqrt(occ_i)  (c = 0: real part, c = 1: imaginary part)
// from Ct = C^T (n_occ x ldct complex). Rows i >= n_occ and columns a >= fs are zero.
template <typename TO>
__global__ void k_build_ct(const double* ct, size_t ldct, const double* occ, size_t n_occ, size_t rp, size_t off0,
                           size_t off1, size_t fs, size_t ld_out, TO* out) {
  const size_t rows = 4 * rp;
  const size_t total = rows * ld_out;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t row = idx / ld_out, a = idx % ld_out;
    const size_t fl = row / (2 * rp), rem = row % (2 * rp), i = rem / 2, cpart = rem % 2;
    double v = 0.0;
    if (i < n_occ && a < fs) {
      const size_t col = (fl == 0 ? off0 : off1) + a;
      const double so = sqrt(occ[i]);
      v = cpart == 0 ? ct[2 * (i * ldct + col)] * so : -ct[2 * (i * ldct + col) + 1] * so;
    }
    out[idx] = from_double<TO>(v);
  }
}

// out = rot(in) over interleaved bf16 (re, im) pairs, rot(re, im) = (im, -re): with it the imaginary
// part of sum_k conj(a_k) b_k is the real dot product of a with rot(b) (count32 = number of pairs,
// a multiple of 4).
__global__ void k_rot_pairs_bf16(const bf16_t* in, bf16_t* out, size_t count32) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4* src = reinterpret_cast<const u4*>(in);
  u4* dst = reinterpret_cast<u4*>(out);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count32 / 4; i += (size_t)gridDim.x * blockDim.x) {
    u4 w = src[i];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = ((w[q] >> 16) | (w[q] << 16)) ^ 0x80000000u;
    dst[i] = w;
  }
}

struct Flavor {
  const Tensor* t;  // tensor of this flavor's spatial type
  size_t off;       // offset of the flavor in the coefficient rows (and in the output)
  size_t fs;        // flavor size
};

// flavors come in pairs sharing a tensor: (0,1), (2,3), ...
template <typename TI>
ExchangeOut exchange_flavors(Context& c, int prec, const std::vector<Flavor>& fl, const double* dCt, size_t ldct,
                             const double* dOcc, size_t n_occ, size_t n_stack) {
  ExchangeOut out;
  const size_t naux = fl[0].t->naux;
  const size_t rp = roundup(std::max<size_t>(n_occ, 1), 8);
  const size_t kdim = naux * rp;          // complex k of the rank-k update
  const size_t ldw = 2 * kdim;            // reals per W row
  TI* W = static_cast<TI*>(wsbuf(c, WS_W, n_stack * ldw * sizeof(TI)));
  // Aux rows past a tensor's last non-zero row add nothing: its half-transform stops there and the
  // rank-k update runs over the common live columns for every row, then over the rest for the rows of
  // the pair with more live rows only (a shared aux index, UkbFockMatrixRi.cpp:46-49, leaves trailing
  // zero rows in the smaller block). W columns past a pair's live rows are never read.
  size_t live[2] = {0, 0}, row0[2] = {0, 0}, rows[2] = {0, 0};
  const int npairs = (int)(fl.size() / 2);
  RR_REQUIRE(npairs >= 1 && npairs <= 2 && fl.size() % 2 == 0, RR_ERR_ARG, "exchange: one or two flavor pairs");
  // stacked row offset of each flavor = its offset (flavors are listed in output order)
  for (size_t pi = 0; pi + 1 < fl.size(); pi += 2) {
    const Flavor& f0 = fl[pi];
    const Flavor& f1 = fl[pi + 1];
    const Tensor& t = *f0.t;
    const size_t nz = live_rows(t);
    live[pi / 2] = nz, row0[pi / 2] = f0.off, rows[pi / 2] = f0.fs + f1.fs;
    TI* lt = static_cast<TI*>(wsbuf(c, WS_LT, 4 * rp * t.ld * sizeof(TI)));
    hipLaunchKernelGGL((k_build_ct<TI>), dim3(grid_for(4 * rp * t.ld)), dim3(256), 0, c.stream, dCt, ldct, dOcc, n_occ, rp,
                       f0.off, f1.off, f0.fs, t.ld, lt);
    RR_HIP(hipGetLastError());
    PhaseTimer pt(c, PH_GEMM1);
    GemmArgs g;
    g.A = tensor_data(t, prec), g.lda = t.ld, g.sAb = t.n * t.ld;
    g.B = lt, g.ldb = t.ld;
    g.C = W + f0.off * ldw, g.C2 = W + f1.off * ldw, g.ncut = (int)(2 * rp);
    g.ldcm = ldw, g.ldcn = 1, g.sCb = 2 * rp;
    g.M = (int)t.n, g.N = (int)(4 * rp), g.K = (int)roundup(t.n, 8), g.nbatch = (int)nz;
    GemmSpec sp;
    sp.prec = prec, sp.out = prec, sp.cfg = gemm1_cfg(c, 4 * rp);
    sp.lds = gemm1_lds(c, t.n, 4 * rp);
    gemm_auto(c, sp, g);
  }
  const int slab = prec == P64 ? 16 : 32;
  const size_t kA = (npairs == 2 ? std::min(live[0], live[1]) : live[0]) * rp;  // complex columns for all rows
  const int big = (npairs == 2 && live[1] > live[0]) ? 1 : 0;
  const size_t kB = npairs == 2 ? std::max(live[0], live[1]) * rp - kA : 0;    // columns for pair `big` only
  auto split_of = [&](size_t m, size_t kd, size_t* kper) {
    const int ns = gemm2_splits(m, 2 * kd, slab);
    *kper = roundup((kd + ns - 1) / ns, slab);  // complex k per split (2*kper reals: multiple of slab)
    return (int)((kd + *kper - 1) / *kper);
  };
  size_t kperA = 0, kperB = 0;
  const int nsA = split_of(n_stack, kA, &kperA);
  const int nsB = kB ? split_of(rows[big], kB, &kperB) : 0;
  const size_t nsq = n_stack * n_stack;
  double* kpart = static_cast<double*>(wsbuf(c, WS_KPART, (size_t)(nsA + nsB) * nsq * 2 * sizeof(double)));
  {
    PhaseTimer pt(c, PH_GEMM2);
    // bf16: the complex update as two real ones over the interleaved rows (the real matrix-core
    // kernels are faster than the complex one): Re K = W . W, Im K = W . rot(W), written into the
    // real and imaginary slots of the complex partial sums (column stride 2)
    constexpr bool kReal = std::is_same<TI, bf16_t>::value;
    const TI* Wrot = nullptr;
    if constexpr (kReal) {
      TI* wr = static_cast<TI*>(wsbuf(c, WS_TMP8 + 500, n_stack * ldw * sizeof(TI)));
      hipLaunchKernelGGL(k_rot_pairs_bf16, dim3(grid_for(n_stack * kdim / 4)), dim3(256), 0, c.stream, W, wr, n_stack * kdim);
      RR_HIP(hipGetLastError());
      Wrot = wr;
    }
    // one part: rows [r0, r0 + m) of W, complex columns [k0, k0 + kd) into partials C (complex)
    auto update = [&](size_t r0, size_t m, size_t k0, size_t kd, size_t kper, int ns, double* C) {
      GemmArgs g;
      g.tri = 1;
      GemmSpec sp;
      sp.prec = prec, sp.out = P64, sp.cfg = gemm2_cfg(c, m);
      if constexpr (kReal) {
        g.lda = ldw, g.ldb = ldw;
        g.ldcm = 2 * n_stack, g.ldcn = 2, g.sCz = 2 * nsq;
        g.M = (int)m, g.N = (int)m, g.K = (int)(2 * kd), g.kper = (int)(2 * kper), g.nsplit = ns;
        sp.lds = gemm2_lds(c, m);
        for (int part = 0; part < 2; ++part) {
          g.A = W + r0 * ldw + 2 * k0;
          g.B = (part ? Wrot : W) + r0 * ldw + 2 * k0;
          g.C = C + part;
          gemm_auto(c, sp, g);
        }
      } else {
        g.A = W + r0 * ldw + 2 * k0, g.lda = kdim, g.B = W + r0 * ldw + 2 * k0, g.ldb = kdim;
        g.C = C, g.ldcm = n_stack, g.ldcn = 1, g.sCz = nsq;
        g.M = (int)m, g.N = (int)m, g.K = (int)kd, g.kper = (int)kper, g.nsplit = ns;
        g.conj = 1;
        sp.cplx = true;
        gemm_auto(c, sp, g);
      }
    };
    update(0, n_stack, 0, kA, kperA, nsA, kpart);
    if (kB) {
      // the other rows of these partial sums are zero
      double* pb = kpart + (size_t)nsA * nsq * 2;
      hipLaunchKernelGGL(k_zero, dim3(grid_for((size_t)nsB * nsq * 2)), dim3(256), 0, c.stream, pb, (size_t)nsB * nsq * 2);
      RR_HIP(hipGetLastError());
      update(row0[big], rows[big], kA, kB, kperB, nsB, pb + 2 * (row0[big] * n_stack + row0[big]));
    }
    out.ktile = 1;
  }
  out.kpart = kpart, out.ksplit = nsA + nsB, out.kstride = nsq;
  return out;
}

ExchangeOut exchange_flavors_prec(Context& c, int prec, const std::vector<Flavor>& fl, const double* dCt, size_t ldct,
                                  const double* dOcc, size_t n_occ, size_t n_stack) {
  if (prec == P64) return exchange_flavors<double>(c, prec, fl, dCt, ldct, dOcc, n_occ, n_stack);
  if (prec == P32) return exchange_flavors<float>(c, prec, fl, dCt, ldct, dOcc, n_occ, n_stack);
  return exchange_flavors<bf16_t>(c, prec, fl, dCt, ldct, dOcc, n_occ, n_stack);
}

// Real density of a complex matrix block sum: out[mu][nu] = Re(P[o0+mu][o0+nu] + P[o1+mu][o1+nu]).
__global__ void k_re_block_sum(const double* p, size_t ldp, size_t o0, size_t o1, size_t n, double* out) {
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n * n; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t mu = idx / n, nu = idx % n;
    out[idx] = p[2 * ((o0 + mu) * ldp + o0 + nu)] + p[2 * ((o1 + mu) * ldp + o1 + nu)];
  }
}

// Coulomb for complex densities over spatial types sharing gamma: returns reduced Jt per type.
// types[k] = tensor, dens[k] = real density (n_k x n_k) for that type.
std::vector<double*> coulomb_shared(Context& c, int prec, const std::vector<const Tensor*>& types,
                                    const std::vector<const double*>& dens) {
  const size_t naux = types[0]->naux;
  double* gamma = static_cast<double*>(wsbuf(c, WS_GAMMA, std::max<size_t>(1, naux) * sizeof(double)));
  for (size_t k = 0; k < types.size(); ++k) {
    const Tensor& t = *types[k];
    const size_t bytes = t.n * t.ld * prec_size(prec);
    void* dt = wsbuf(c, WS_DT + 1000 * (int)k, bytes);
    if (prec == P64)
      hipLaunchKernelGGL((k_pack<double>), dim3(grid_for(t.n * t.ld)), dim3(256), 0, c.stream, dens[k], t.n, 1, t.n, t.n,
                         t.n, t.ld, 1.0, (double*)dt);
    else if (prec == P32)
      hipLaunchKernelGGL((k_pack<float>), dim3(grid_for(t.n * t.ld)), dim3(256), 0, c.stream, dens[k], t.n, 1, t.n, t.n,
                         t.n, t.ld, 1.0, (float*)dt);
    else
      hipLaunchKernelGGL((k_pack<bf16_t>), dim3(grid_for(t.n * t.ld)), dim3(256), 0, c.stream, dens[k], t.n, 1, t.n, t.n,
                         t.n, t.ld, 1.0, (bf16_t*)dt);
    RR_HIP(hipGetLastError());
    if (prec == P64) gamma_stream<double>(c, t, (const double*)dt, gamma, k > 0);
    else if (prec == P32) gamma_stream<float>(c, t, (const float*)dt, gamma, k > 0);
    else gamma_stream<bf16_t>(c, t, (const bf16_t*)dt, gamma, k > 0);
  }
  std::vector<double*> jt;
  for (size_t k = 0; k < types.size(); ++k) {
    const Tensor& t = *types[k];
    double* jpart = nullptr;
    double* jp_buf = static_cast<double*>(wsbuf(c, WS_JPART + 1000 * (int)(k + 1), 32 * t.n * t.ld * sizeof(double)));
    int js;
    if (prec == P64) js = coulomb_stream<double>(c, t, gamma, jp_buf, &jpart);
    else if (prec == P32) js = coulomb_stream<float>(c, t, gamma, jp_buf, &jpart);
    else js = coulomb_stream<bf16_t>(c, t, gamma, jp_buf, &jpart);
    double* red = static_cast<double*>(wsbuf(c, WS_OUT_J + 1000 * (int)(k + 1), t.n * t.ld * sizeof(double)));
    hipLaunchKernelGGL(k_reduce_parts, dim3(grid_for(t.n * t.ld)), dim3(256), 0, c.stream, jpart, js, t.n * t.ld,
                       t.n * t.ld, red);
    RR_HIP(hipGetLastError());
    jt.push_back(red);
  }
  return jt;
}

// G = J - K (plus h) in a complex nd x nd output. Coulomb blocks: (off, size, type index).
void cplx_assemble(Context& c, const ExchangeOut* ex, size_t nd, const double* dh, double* dout,
                   const std::vector<std::pair<size_t, size_t>>& blocks, const std::vector<int>& btype,
                   const std::vector<double*>& jt, const std::vector<const Tensor*>& types) {
  PhaseTimer pt(c, PH_ASM);
  CplxAsm a;
  a.nd = nd;
  if (ex && ex->kpart) a.kpart = ex->kpart, a.ksplit = ex->ksplit, a.kstride = ex->kstride, a.ktile = ex->ktile;
  a.kfac = -1.0, a.h = dh, a.out = dout;
  a.nblk = (int)blocks.size();
  for (size_t b = 0; b < blocks.size(); ++b) {
    a.off[b] = blocks[b].first, a.sz[b] = blocks[b].second;
    a.jsrc[b] = jt[btype[b]], a.jld[b] = types[btype[b]]->ld;
  }
  hipLaunchKernelGGL(k_cplx_asm, dim3(grid_for(nd * nd)), dim3(256), 0, c.stream, a);
  RR_HIP(hipGetLastError());
}

__global__ void k_copy_real(const double* in, double* out, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) out[i] = in[i];
}

// Core of riFockTwoElectronDirect on device arrays: G (n_ukb x n_ukb complex) = J - K.
void two_electron_direct(Context& c, int prec, const Tensor& tl, const Tensor& ts, size_t nl, size_t ns,
                         const double* dP, const double* dCt, size_t ldct, const double* dOcc, size_t n_occ, double* dG) {
  const size_t n = 2 * nl + 2 * ns;
  double* dl = static_cast<double*>(wsbuf(c, WS_DENS2, std::max<size_t>(1, nl * nl + ns * ns) * sizeof(double)));
  double* ds = dl + nl * nl;
  if (nl) hipLaunchKernelGGL(k_re_block_sum, dim3(grid_for(nl * nl)), dim3(256), 0, c.stream, dP, n, 0, nl, nl, dl);
  if (ns) hipLaunchKernelGGL(k_re_block_sum, dim3(grid_for(ns * ns)), dim3(256), 0, c.stream, dP, n, 2 * nl, 2 * nl + ns, ns, ds);
  RR_HIP(hipGetLastError());
  std::vector<const Tensor*> types = {&tl, &ts};
  std::vector<const double*> dens = {dl, ds};
  std::vector<double*> jt = coulomb_shared(c, prec, types, dens);
  ExchangeOut ex;
  const bool have_k = n_occ > 0;
  if (have_k) {
    std::vector<Flavor> fl = {{&tl, 0, nl}, {&tl, nl, nl}, {&ts, 2 * nl, ns}, {&ts, 2 * nl + ns, ns}};
    ex = exchange_flavors_prec(c, prec, fl, dCt, ldct, dOcc, n_occ, n);
  }
  cplx_assemble(c, have_k ? &ex : nullptr, n, nullptr, dG, {{0, nl}, {nl, nl}, {2 * nl, ns}, {2 * nl + ns, ns}},
                {0, 0, 1, 1}, jt, types);
}

// ================================================================================================
// MO transforms
// ================================================================================================
// Real two-sided transform of resident tensor rows: out_P = C^T B_P C (C n x nmo), P in [0, naux).
// step 1: S1_P[q][mu'] = sum_nu Ct[q][nu] Bt_P[mu'][nu]  (= (C^T B_P)[q][mu'])
// step 2: out_P[p][q]  = sum_mu' S1_P[p][mu'] Ct[q][mu']
template <typename TI, typename TO>
void mo_real_slot(Context& c, const Tensor& t, int prec, const double* dC, size_t nmo, TO* dout) {
  const size_t n = t.n, ld = t.ld;
  TI* ct = static_cast<TI*>(wsbuf(c, WS_MO_A, std::max<size_t>(1, nmo) * ld * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(nmo * ld)), dim3(256), 0, c.stream, dC, nmo, 1, nmo, nmo, n, ld, 1.0, ct);
  RR_HIP(hipGetLastError());
  const size_t per = nmo * ld;
  size_t chunk = std::max<size_t>(1, std::min<size_t>(t.naux, (size_t(512) << 20) / std::max<size_t>(1, per * sizeof(TI))));
  TI* s1 = static_cast<TI*>(wsbuf(c, WS_MO_S1, chunk * per * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(chunk * per)), dim3(256), 0, c.stream, (const double*)nullptr, (size_t)0, 0,
                     chunk * nmo, (size_t)0, (size_t)0, ld, 0.0, s1);  // zero (pads stay zero)
  RR_HIP(hipGetLastError());
  const int op = std::is_same<TI, double>::value ? P64 : (std::is_same<TI, float>::value ? P32 : P16);
  const int oo = std::is_same<TO, double>::value ? P64 : P32;
  for (size_t p0 = 0; p0 < t.naux; p0 += chunk) {
    const size_t rows = std::min(chunk, t.naux - p0);
    {
      PhaseTimer pt(c, PH_GEMM1);
      GemmArgs g;
      g.A = ct, g.lda = ld;
      g.B = static_cast<const TI*>(tensor_data(t, prec)) + p0 * n * ld, g.ldb = ld, g.sBb = n * ld;
      g.C = s1, g.ldcm = ld, g.ldcn = 1, g.sCb = per;
      g.M = (int)nmo, g.N = (int)n, g.K = (int)roundup(n, 8), g.nbatch = (int)rows;
      GemmSpec sp;
      sp.prec = op, sp.out = op, sp.cfg = 2;
      gemm_auto(c, sp, g);
    }
    {
      PhaseTimer pt(c, PH_GEMM2);
      GemmArgs g;
      g.A = s1, g.lda = ld, g.sAb = per;
      g.B = ct, g.ldb = ld;
      g.C = dout + p0 * nmo * nmo, g.ldcm = nmo, g.ldcn = 1, g.sCb = nmo * nmo;
      g.M = (int)nmo, g.N = (int)nmo, g.K = (int)ld, g.nbatch = (int)rows;
      GemmSpec sp;
      sp.prec = op, sp.out = oo, sp.cfg = 2;
      gemm_auto(c, sp, g);
    }
  }
}

// Step-1 operand of the complex MO transform for one block: rows m = c*nt + p hold
//   c = 0: Re C_b[mu][p],  c = 1: -Im C_b[mu][p]   (conj(C_b)^T split into real rows), mu < fs.
template <typename TO>
__global__ void k_mo_a1(const double* cu, size_t nt, size_t off, size_t fs, size_t ld, TO* out) {
  const size_t total = 2 * nt * ld;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t m = idx / ld, mu = idx % ld, cp = m / nt, pp = m % nt;
    double v = 0.0;
    if (mu < fs) v = cp == 0 ? cu[2 * ((off + mu) * nt + pp)] : -cu[2 * ((off + mu) * nt + pp) + 1];
    out[idx] = from_double<TO>(v);
  }
}
// Step-2 operand: row (2q + d), concatenated over blocks b and parts c (k' = 2 off_b + c fs_b + nu):
//   d = 0: c = 0 -> Re C_b[nu][q], c = 1 -> -Im C_b[nu][q];  d = 1: c = 0 -> Im, c = 1 -> Re.
template <typename TO>
__global__ void k_mo_b2(const double* cu, size_t nt, size_t nukb, size_t ldk, TO* out) {
  const size_t total = 2 * nt * ldk;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t row = idx / ldk, k = idx % ldk, q = row / 2, d = row % 2;
    double v = 0.0;
    if (k < 2 * nukb) {
      // locate block: rows of C are ordered by block, k' = 2*off + c*fs + nu; we pass the block
      // table through the row index: handled by the caller writing one block at a time (see below)
      v = 0.0;
    }
    out[idx] = from_double<TO>(v);
  }
}
template <typename TO>
__global__ void k_mo_b2_block(const double* cu, size_t nt, size_t off, size_t fs, size_t ldk, TO* out) {
  // fills k' in [2 off, 2 off + 2 fs) of every row
  const size_t total = 2 * nt * 2 * fs;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t row = idx / (2 * fs), kk = idx % (2 * fs), q = row / 2, d = row % 2, cp = kk / fs, nu = kk % fs;
    const double re = cu[2 * ((off + nu) * nt + q)], im = cu[2 * ((off + nu) * nt + q) + 1];
    double v;
    if (d == 0) v = cp == 0 ? re : -im;
    else v = cp == 0 ? im : re;
    out[row * ldk + 2 * off + kk] = from_double<TO>(v);
  }
}

// Complex MO transform of resident real tensors (buildRiMoEri / aoCholeskyToMoSpinor):
// out_P[p][q] = sum_b sum_{mu,nu in b} conj(C_b[mu][p]) B^b_P[mu][nu] C_b[nu][q]; out is complex TO.
template <typename TI, typename TO>
void mo_cplx_slot(Context& c, const std::vector<Flavor>& fl, int prec, const double* dC, size_t nt, size_t nukb, TO* dout) {
  const size_t naux = fl[0].t->naux;
  const size_t ldk = roundup(2 * nukb, 8);
  const int op = std::is_same<TI, double>::value ? P64 : (std::is_same<TI, float>::value ? P32 : P16);
  const int oo = std::is_same<TO, double>::value ? P64 : P32;
  // step-2 operand (shared over P)
  TI* b2 = static_cast<TI*>(wsbuf(c, WS_MO_B, 2 * nt * ldk * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(2 * nt * ldk)), dim3(256), 0, c.stream, (const double*)nullptr, (size_t)0, 0,
                     2 * nt, (size_t)0, (size_t)0, ldk, 0.0, b2);
  for (const Flavor& f : fl)
    if (f.fs) hipLaunchKernelGGL((k_mo_b2_block<TI>), dim3(grid_for(4 * nt * f.fs)), dim3(256), 0, c.stream, dC, nt, f.off, f.fs,
                                 ldk, b2);
  RR_HIP(hipGetLastError());
  const size_t per = nt * ldk;  // S1 per P (rows p, k' of length ldk)
  size_t chunk = std::max<size_t>(1, std::min<size_t>(naux, (size_t(768) << 20) / std::max<size_t>(1, per * sizeof(TI))));
  TI* s1 = static_cast<TI*>(wsbuf(c, WS_MO_S1, chunk * per * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(chunk * per)), dim3(256), 0, c.stream, (const double*)nullptr, (size_t)0, 0,
                     chunk * nt, (size_t)0, (size_t)0, ldk, 0.0, s1);
  RR_HIP(hipGetLastError());
  for (size_t p0 = 0; p0 < naux; p0 += chunk) {
    const size_t rows = std::min(chunk, naux - p0);
    for (const Flavor& f : fl) {
      if (!f.fs) continue;
      const Tensor& t = *f.t;
      TI* a1 = static_cast<TI*>(wsbuf(c, WS_MO_A, 2 * nt * t.ld * sizeof(TI)));
      hipLaunchKernelGGL((k_mo_a1<TI>), dim3(grid_for(2 * nt * t.ld)), dim3(256), 0, c.stream, dC, nt, f.off, f.fs, t.ld, a1);
      RR_HIP(hipGetLastError());
      PhaseTimer pt(c, PH_GEMM1);
      GemmArgs g;
      g.A = a1, g.lda = t.ld;
      g.B = static_cast<const TI*>(tensor_data(t, prec)) + p0 * t.n * t.ld, g.ldb = t.ld, g.sBb = t.n * t.ld;
      g.C = s1 + 2 * f.off, g.C2 = s1 + 2 * f.off + f.fs, g.mcut = (int)nt;
      g.ldcm = ldk, g.ldcn = 1, g.sCb = per;
      g.M = (int)(2 * nt), g.N = (int)t.n, g.K = (int)roundup(t.n, 8), g.nbatch = (int)rows;
      GemmSpec sp;
      sp.prec = op, sp.out = op, sp.cfg = 2;
      gemm_auto(c, sp, g);
    }
    PhaseTimer pt(c, PH_GEMM2);
    GemmArgs g;
    g.A = s1, g.lda = ldk, g.sAb = per;
    g.B = b2, g.ldb = ldk;
    g.C = dout + p0 * nt * nt * 2, g.ldcm = 2 * nt, g.ldcn = 1, g.sCb = 2 * nt * nt;
    g.M = (int)nt, g.N = (int)(2 * nt), g.K = (int)ldk, g.nbatch = (int)rows;
    GemmSpec sp;
    sp.prec = op, sp.out = oo, sp.cfg = 2;
    gemm_auto(c, sp, g);
  }
}

// Pack complex n x n matrices (batch) into rows padded to ld complex, converting to TI.
template <typename TO>
__global__ void k_pack_cplx(const double* in, size_t nb, size_t n, size_t ld, TO* out) {
  const size_t total = nb * n * ld;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t b = idx / (n * ld), rem = idx % (n * ld), r = rem / ld, k = rem % ld;
    double re = 0.0, im = 0.0;
    if (k < n) re = in[2 * ((b * n + r) * n + k)], im = in[2 * ((b * n + r) * n + k) + 1];
    out[2 * idx] = from_double<TO>(re);
    out[2 * idx + 1] = from_double<TO>(im);
  }
}
// Transposed complex pack: out[r][k] = in[k][r] (one n x n matrix), padded to ld.
template <typename TO>
__global__ void k_pack_cplx_t(const double* in, size_t n, size_t ld, TO* out) {
  const size_t total = n * ld;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t r = idx / ld, k = idx % ld;
    double re = 0.0, im = 0.0;
    if (k < n) re = in[2 * (k * n + r)], im = in[2 * (k * n + r) + 1];
    out[2 * idx] = from_double<TO>(re);
    out[2 * idx + 1] = from_double<TO>(im);
  }
}

// Complex rotation B'_P = U^H B_P U of an input tensor b (naux, n*n complex, device):
// step 1: St_P[q][a] = sum_b Ut[q][b] B_P[a][b];  step 2: out_P[p][q] = sum_a conj(Ut[p][a]) St_P[q][a].
template <typename TI>
void rotate_cplx(Context& c, int prec, size_t n, size_t naux, const double* db, const double* du, double* dout) {
  const size_t ld = roundup(std::max<size_t>(n, 1), 8);
  TI* ut = static_cast<TI*>(wsbuf(c, WS_MO_A, 2 * n * ld * sizeof(TI)));
  hipLaunchKernelGGL((k_pack_cplx_t<TI>), dim3(grid_for(n * ld)), dim3(256), 0, c.stream, du, n, ld, ut);
  const size_t per = n * ld * 2;  // reals per packed matrix
  size_t chunk = std::max<size_t>(1, std::min<size_t>(naux, (size_t(512) << 20) / std::max<size_t>(1, per * sizeof(TI))));
  TI* bp = static_cast<TI*>(wsbuf(c, WS_MO_B, chunk * per * sizeof(TI)));
  TI* st = static_cast<TI*>(wsbuf(c, WS_MO_S1, chunk * per * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(chunk * per)), dim3(256), 0, c.stream, (const double*)nullptr, (size_t)0, 0,
                     chunk * n, (size_t)0, (size_t)0, 2 * ld, 0.0, st);
  RR_HIP(hipGetLastError());
  const int op = std::is_same<TI, double>::value ? P64 : (std::is_same<TI, float>::value ? P32 : P16);
  for (size_t p0 = 0; p0 < naux; p0 += chunk) {
    const size_t rows = std::min(chunk, naux - p0);
    hipLaunchKernelGGL((k_pack_cplx<TI>), dim3(grid_for(rows * n * ld)), dim3(256), 0, c.stream, db + p0 * n * n * 2, rows, n,
                       ld, bp);
    RR_HIP(hipGetLastError());
    {
      PhaseTimer pt(c, PH_GEMM1);
      GemmArgs g;
      g.A = ut, g.lda = ld;
      g.B = bp, g.ldb = ld, g.sBb = n * ld;
      g.C = st, g.ldcm = ld, g.ldcn = 1, g.sCb = n * ld;
      g.M = (int)n, g.N = (int)n, g.K = (int)ld, g.nbatch = (int)rows;
      GemmSpec sp;
      sp.prec = op, sp.out = op, sp.cplx = true, sp.cfg = 2;
      gemm_auto(c, sp, g);
    }
    PhaseTimer pt(c, PH_GEMM2);
    GemmArgs g;
    g.A = ut, g.lda = ld;
    g.B = st, g.ldb = ld, g.sBb = n * ld;
    g.C = dout + p0 * n * n * 2, g.ldcm = n, g.ldcn = 1, g.sCb = n * n;
    g.M = (int)n, g.N = (int)n, g.K = (int)ld, g.nbatch = (int)rows, g.conj = 1;
    GemmSpec sp;
    sp.prec = op, sp.out = P64, sp.cplx = true, sp.cfg = 2;
    gemm_auto(c, sp, g);
  }
}

// Real rotation out_P = U^T X_P U of an input tensor x (naux, n*n real, device):
// step 1: S_P[q][a] = sum_b Ut[q][b] X_P[a][b];  step 2: out_P[p][q] = sum_a Ut[p][a] S_P[q][a].
template <typename TI>
void rotate_real(Context& c, int prec, size_t n, size_t naux, const double* dx, const double* du_s, double* dout) {
  const size_t ld = roundup(std::max<size_t>(n, 1), 8);
  TI* ut = static_cast<TI*>(wsbuf(c, WS_MO_A, n * ld * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(n * ld)), dim3(256), 0, c.stream, du_s, n, 1, n, n, n, ld, 1.0, ut);
  const size_t per = n * ld;
  size_t chunk = std::max<size_t>(1, std::min<size_t>(naux, (size_t(512) << 20) / std::max<size_t>(1, per * sizeof(TI))));
  TI* xp = static_cast<TI*>(wsbuf(c, WS_MO_B, chunk * per * sizeof(TI)));
  TI* st = static_cast<TI*>(wsbuf(c, WS_MO_S1, chunk * per * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(chunk * per)), dim3(256), 0, c.stream, (const double*)nullptr, (size_t)0, 0,
                     chunk * n, (size_t)0, (size_t)0, ld, 0.0, st);
  RR_HIP(hipGetLastError());
  const int op = std::is_same<TI, double>::value ? P64 : (std::is_same<TI, float>::value ? P32 : P16);
  for (size_t p0 = 0; p0 < naux; p0 += chunk) {
    const size_t rows = std::min(chunk, naux - p0);
    hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(rows * n * ld)), dim3(256), 0, c.stream, dx + p0 * n * n, n, 0, rows * n,
                       rows * n, n, ld, 1.0, xp);
    RR_HIP(hipGetLastError());
    {
      PhaseTimer pt(c, PH_GEMM1);
      GemmArgs g;
      g.A = ut, g.lda = ld;
      g.B = xp, g.ldb = ld, g.sBb = n * ld;
      g.C = st, g.ldcm = ld, g.ldcn = 1, g.sCb = n * ld;
      g.M = (int)n, g.N = (int)n, g.K = (int)ld, g.nbatch = (int)rows;
      GemmSpec sp;
      sp.prec = op, sp.out = op, sp.cfg = 2;
      gemm_auto(c, sp, g);
    }
    PhaseTimer pt(c, PH_GEMM2);
    GemmArgs g;
    g.A = ut, g.lda = ld;
    g.B = st, g.ldb = ld, g.sBb = n * ld;
    g.C = dout + p0 * n * n, g.ldcm = n, g.ldcn = 1, g.sCb = n * n;
    g.M = (int)n, g.N = (int)n, g.K = (int)ld, g.nbatch = (int)rows;
    GemmSpec sp;
    sp.prec = op, sp.out = P64, sp.cfg = 2;
    gemm_auto(c, sp, g);
  }
}

__global__ void k_positive_block(const double* b, size_t nt, size_t naux, size_t off, double* out) {
  const size_t m = nt - off;
  const size_t total = naux * m * m;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t P = idx / (m * m), rem = idx % (m * m), a = rem / m, cc = rem % m;
    const size_t src = (P * nt + (a + off)) * nt + (cc + off);
    out[2 * idx] = b[2 * src];
    out[2 * idx + 1] = b[2 * src + 1];
  }
}

__global__ void k_embedded_b(const double* b, size_t ns, size_t naux, double* out) {
  const size_t n = 2 * ns;
  const size_t total = naux * n * n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t P = idx / (n * n), rem = idx % (n * n), r = rem / n, cc = rem % n;
    double v = 0.0;
    if ((r < ns) == (cc < ns)) v = b[(P * ns + (r % ns)) * ns + (cc % ns)];
    out[idx] = v;
  }
}

// ================================================================================================
// Cholesky D-form exchange: K = sum_L B_L P B_L for symmetric vectors (AoCholesky.cpp:133-136)
// step 1: Z_L[b][mu] = sum_lambda Bt_L[b][lambda] Pop[mu][lambda]     (batched over L)
// step 2: K[a][b]    = sum_{L,mu} B_L[a][mu] Z_L[b][mu]                 (k-batched over L)
// For a complex P block, Pop holds the real rows of P and then its imaginary rows, so Z_L rows
// are [Re | Im] and step 2 writes interleaved complex output.
// ================================================================================================
template <typename TI>
void cholesky_exchange(Context& c, const Tensor& t, int prec, const double* dPop, size_t prow, double* kout, size_t ldk,
                       bool cplx_rows) {
  const size_t n = t.n, ld = t.ld, nchol = t.naux;
  const size_t zw = cplx_rows ? 2 * ld : ld;  // Z row length
  TI* pop = static_cast<TI*>(wsbuf(c, WS_MO_A, prow * ld * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(prow * ld)), dim3(256), 0, c.stream, dPop, n, 0, prow, prow, n, ld, 1.0, pop);
  RR_HIP(hipGetLastError());
  TI* z = static_cast<TI*>(wsbuf(c, WS_W, std::max<size_t>(1, nchol) * n * zw * sizeof(TI)));
  hipLaunchKernelGGL((k_pack<TI>), dim3(grid_for(nchol * n * zw)), dim3(256), 0, c.stream, (const double*)nullptr, (size_t)0, 0,
                     nchol * n, (size_t)0, (size_t)0, zw, 0.0, z);
  RR_HIP(hipGetLastError());
  const int op = prec;
  {
    PhaseTimer pt(c, PH_GEMM1);
    GemmArgs g;
    g.A = tensor_data(t, prec), g.lda = ld, g.sAb = n * ld;
    g.B = pop, g.ldb = ld;
    g.C = z, g.ldcm = zw, g.ldcn = 1, g.sCb = n * zw;
    if (cplx_rows) g.C2 = z + ld, g.ncut = (int)n;
    g.M = (int)n, g.N = (int)prow, g.K = (int)ld, g.nbatch = (int)nchol;
    GemmSpec sp;
    sp.prec = op, sp.out = op, sp.cfg = 2;
    gemm_auto(c, sp, g);
  }
  PhaseTimer pt(c, PH_GEMM2);
  GemmArgs g;
  g.A = tensor_data(t, prec), g.lda = ld, g.sAk = n * ld;
  g.B = z, g.ldb = ld, g.sBk = n * zw;
  g.C = kout, g.ldcm = ldk, g.ldcn = 1;
  g.M = (int)n, g.N = (int)(cplx_rows ? 2 * n : n), g.K = (int)ld, g.kbatch = (int)nchol;
  GemmSpec sp;
  sp.prec = op, sp.out = P64, sp.cfg = 2;
  // The output alone gives few 128 x 128 tiles (16 at n = 500) for a contraction over n_chol * ld:
  // split the flat k into up to 2048 workgroups' worth of partial sums (at most 512 MB of them),
  // summed in split order afterwards. Each split covers whole k-slabs (kper a multiple of the slab).
  const size_t tiles = ((n + 127) / 128) * (((size_t)g.N + 127) / 128);
  const size_t ktot = nchol * ld, slab = op == P64 ? 16 : 32, outsz = n * (size_t)g.N;
  const size_t ns = std::min<size_t>({256, std::max<size_t>(1, 2048 / tiles), std::max<size_t>(1, ktot / 4096),
                                      std::max<size_t>(1, (size_t(512) << 20) / (outsz * sizeof(double)))});
  if (ns > 1 && ldk == (size_t)g.N) {
    const size_t kper = roundup((ktot + ns - 1) / ns, slab);
    const int nsplit = (int)((ktot + kper - 1) / kper);
    double* part = static_cast<double*>(wsbuf(c, WS_KPART, (size_t)nsplit * outsz * sizeof(double)));
    g.C = part, g.sCz = (long long)outsz, g.kper = (int)kper, g.nsplit = nsplit;
    gemm_auto(c, sp, g);
    hipLaunchKernelGGL(k_reduce_parts, dim3(grid_for(outsz)), dim3(256), 0, c.stream, part, nsplit, outsz, outsz, kout);
    RR_HIP(hipGetLastError());
    return;
  }
  gemm_auto(c, sp, g);
}

void cholesky_exchange_prec(Context& c, const Tensor& t, int prec, const double* dPop, size_t prow, double* kout, size_t ldk,
                            bool cplx_rows) {
  if (prec == P64) cholesky_exchange<double>(c, t, prec, dPop, prow, kout, ldk, cplx_rows);
  else if (prec == P32) cholesky_exchange<float>(c, t, prec, dPop, prow, kout, ldk, cplx_rows);
  else cholesky_exchange<bf16_t>(c, t, prec, dPop, prow, kout, ldk, cplx_rows);
}

// F = h + J - 1/2 K (real) with J from Jt partials (transposed) and K given (n x n)
__global__ void k_fock_real_k(size_t n, size_t ld, const double* jpart, int jsplit, size_t jstride, const double* k,
                              const double* h, double* f) {
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n * n; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t mu = idx / n, nu = idx % n;
    double j = 0.0;
    for (int s = 0; s < jsplit; ++s) j += jpart[s * jstride + nu * ld + mu];
    f[idx] = h[idx] + j - 0.5 * k[idx];
  }
}
// X2C Cholesky assembly: F = h + blockdiag(J, J) - [K_su] (K_su complex n x n, layout [s][u])
__global__ void k_fock_x2c_chol(size_t n, size_t ld, const double* jt, const double* ksu, const double* h, double* f) {
  const size_t n2 = 2 * n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n2 * n2; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t r = idx / n2, cc = idx % n2, s = r / n, u = cc / n, a = r % n, b = cc % n;
    double re = h[2 * idx], im = h[2 * idx + 1];
    if (s == u) re += jt[b * ld + a];
    const double* k = ksu + (s * 2 + u) * n * n * 2;
    re -= k[2 * (a * n + b)];
    im -= k[2 * (a * n + b) + 1];
    f[2 * idx] = re;
    f[2 * idx + 1] = im;
  }
}
// Split a complex block P_su (rows of a 2n x 2n complex matrix) into Pop rows [Re rows ; Im rows]
__global__ void k_split_block(const double* p, size_t n, size_t s, size_t u, double* out) {
  const size_t n2 = 2 * n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < 2 * n * n; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t part = idx / (n * n), rem = idx % (n * n), a = rem / n, b = rem % n;
    out[idx] = p[2 * ((s * n + a) * n2 + u * n + b) + part];
  }
}

// FP64 MO tensors written straight into device memory, for the fused functional Fock builds
// (riJkOnlyFockMatrix / riPnofFockMatrix and their NON_REL counterparts).
void mo_tensor_rkb(Context& c, int slot_ll, int slot_ss, size_t nl, size_t ns, const double* dC_ukb, size_t nt,
                   double* dout) {
  Tensor& tl = get_slot(c, slot_ll);
  RR_REQUIRE(tl.n == nl, RR_ERR_DIM, "buildRiMoEri: c_ukb row count does not match 2*nLarge+2*nSmall");
  RR_REQUIRE(tl.f64, RR_ERR_ARG, "tensor slot has no copy in the requested precision");
  std::vector<Flavor> fl = {{&tl, 0, nl}, {&tl, nl, nl}};
  if (ns > 0) {
    Tensor& ts = get_slot(c, slot_ss);
    RR_REQUIRE(ts.n == ns && ts.naux == tl.naux, RR_ERR_DIM, "buildRiMoEri: c_ukb row count does not match 2*nLarge+2*nSmall");
    RR_REQUIRE(ts.f64, RR_ERR_ARG, "tensor slot has no copy in the requested precision");
    fl.push_back({&ts, 2 * nl, ns});
    fl.push_back({&ts, 2 * nl + ns, ns});
  }
  mo_cplx_slot<double, double>(c, fl, P64, dC_ukb, nt, 2 * nl + 2 * ns, dout);
}

void mo_tensor_nonrel_embedded(Context& c, int slot, size_t n, const double* dC_spatial, double* dout) {
  Tensor& t = get_slot(c, slot);
  RR_REQUIRE(t.n == n, RR_ERR_DIM, "buildRiNonRelSpinMoEri: eri3_L column count does not match n_spatial*n_spatial");
  RR_REQUIRE(t.f64, RR_ERR_ARG, "tensor slot has no copy in the requested precision");
  double* mo = static_cast<double*>(wsbuf(c, WS_MO_OUT, std::max<size_t>(1, t.naux * n * n) * sizeof(double)));
  mo_real_slot<double, double>(c, t, P64, dC_spatial, n, mo);
  hipLaunchKernelGGL(k_embedded_b, dim3(grid_for(t.naux * 4 * n * n)), dim3(256), 0, c.stream, mo, n, t.naux, dout);
  RR_HIP(hipGetLastError());
}

size_t slot_naux(Context& c, int slot) { return get_slot(c, slot).naux; }

bool set_phase_timing(bool on) {
  const bool was = PhaseTimer::enabled;
  PhaseTimer::enabled = on;
  return was;
}

// Allocates t (freeing what it held) and stores eri3 (n_aux, n*n), host or device, in the slot layout.
void tensor_free(Tensor& t) {
  if (t.f64) (void)hipFree(t.f64);
  if (t.f32) (void)hipFree(t.f32);
  if (t.b16) (void)hipFree(t.b16);
  t = Tensor();
}

// flag[0] = 1 if some row block of src (rows x n x n) is not exactly symmetric; flag[1] = max(flag[1],
// 1 + index (offset by p0) of the last row holding a non-zero value).
__global__ void k_sym_check(const double* src, size_t rows, size_t n, size_t p0, int* flag) {
  __shared__ int last[256];
  const size_t nn = n * n, total = rows * nn;
  int mine = 0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t P = idx / nn, rem = idx % nn, mu = rem / n, nu = rem % n;
    const double v = src[idx];
    if (nu > mu && !(v == src[P * nn + nu * n + mu])) flag[0] = 1;
    if (v != 0.0) mine = max(mine, (int)(p0 + P + 1));
  }
  last[threadIdx.x] = mine;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) last[threadIdx.x] = max(last[threadIdx.x], last[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && last[0] > 0) atomicMax(flag + 1, last[0]);
}

void tensor_store(Context& c, Tensor& t, const double* eri3, size_t naux, size_t n, unsigned copies) {
  RR_REQUIRE(copies & 7u, RR_ERR_ARG, "no tensor copy requested");
  RR_REQUIRE(eri3 || naux * n == 0, RR_ERR_ARG, "null tensor");
  RR_HIP(hipStreamSynchronize(c.stream));
  tensor_free(t);
  t.naux = naux, t.n = n, t.ld = tensor_ld(n);
  const size_t cnt = std::max<size_t>(1, naux * n * t.ld);
  if (copies & RR_COPY_FP64) RR_HIP(hipMalloc(&t.f64, cnt * sizeof(double)));
  if (copies & RR_COPY_FP32) RR_HIP(hipMalloc(&t.f32, cnt * sizeof(float)));
  if (copies & RR_COPY_BF16) RR_HIP(hipMalloc(&t.b16, cnt * sizeof(bf16_t)));
  if (naux == 0 || n == 0) return;
  const bool dev = is_device_ptr(eri3);
  const size_t rowbytes = n * n * sizeof(double);
  if (dev) RR_VALIDATE_EXTENT(eri3, naux * rowbytes, "3-index tensor source");
  size_t chunk = std::max<size_t>(1, std::min<size_t>(naux, (size_t(256) << 20) / rowbytes));
  chunk = std::min<size_t>(chunk, 65535);
  double* stage = dev ? nullptr : static_cast<double*>(wsbuf(c, WS_STAGE, chunk * rowbytes));
  // Exactly symmetric blocks (mu nu|P) = (nu mu|P) let the Coulomb stream read only the upper triangle.
  int* asym = static_cast<int*>(wsbuf(c, WS_TMP9 + 77, 2 * sizeof(int)));
  RR_HIP(hipMemsetAsync(asym, 0, 2 * sizeof(int), c.stream));
  for (size_t p0 = 0; p0 < naux; p0 += chunk) {
    const size_t rows = std::min(chunk, naux - p0);
    const double* src = eri3 + p0 * n * n;
    if (!dev) {
      RR_HIP(hipMemcpyAsync(stage, src, rows * rowbytes, hipMemcpyHostToDevice, c.stream));
      src = stage;
    }
    dim3 grid((unsigned)((t.ld + 31) / 32), (unsigned)((n + 31) / 32), (unsigned)rows);  // every padding column
    if (t.f64) hipLaunchKernelGGL(k_store_rows<double>, grid, dim3(256), 0, c.stream, src, n, t.ld, t.f64 + p0 * n * t.ld);
    if (t.f32) hipLaunchKernelGGL(k_store_rows<float>, grid, dim3(256), 0, c.stream, src, n, t.ld, t.f32 + p0 * n * t.ld);
    if (t.b16) hipLaunchKernelGGL(k_store_rows<bf16_t>, grid, dim3(256), 0, c.stream, src, n, t.ld, t.b16 + p0 * n * t.ld);
    hipLaunchKernelGGL(k_sym_check, dim3(grid_for(rows * n * n)), dim3(256), 0, c.stream, src, rows, n, p0, asym);
    RR_HIP(hipGetLastError());
    if (!dev) RR_HIP(hipStreamSynchronize(c.stream));
  }
  int h_asym[2] = {0, 0};
  RR_HIP(hipMemcpyAsync(h_asym, asym, sizeof(h_asym), hipMemcpyDeviceToHost, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
#ifdef RR_VALIDATE
  h_asym[1] = (int)naux;  // no kernel ran: every row is taken as non-zero unless the replay says otherwise
#endif
  RR_CONTROL("upload_asymmetric", h_asym[0]);
  RR_CONTROL("upload_nonzero_rows", h_asym[1]);
  t.symmetric = h_asym[0] == 0;
  t.naux_nz = (size_t)h_asym[1];
}

// ================================================================================================
// Grouped launches: one launch runs a kernel body over many independent problems (segments), each
// with its own arguments and block count; block b runs segment s (first[s] <= b < first[s+1]) as its
// block b - first[s]. Used to build a whole batch of molecules in a few launches. Every segment
// gets the grid the single-problem launch would use, so results are identical to those launches.
// ================================================================================================
template <typename Op>
__global__ __launch_bounds__(256) void k_grouped(const typename Op::Args* args, const unsigned* first, int nseg) {
  const unsigned b = blockIdx.x;
  int lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid;
  }
  const typename Op::Args a = args[lo];  // uniform: held in scalar registers
  Op::run(a, b - first[lo]);
}

template <typename TO>
struct OpPack {
  struct Args {
    const double* in;
    size_t ld_in;
    int trans;
    size_t R, Rv, Cc, ld_out;
    double scale;
    TO* out;
    unsigned gx;
  };
  __device__ static void run(const Args& a, unsigned lb) {
    k_pack_body<TO>(a.in, a.ld_in, a.trans, a.R, a.Rv, a.Cc, a.ld_out, a.scale, a.out, lb, 0, a.gx);
  }
};
struct OpDfac {
  struct Args {
    const double* f0;
    size_t r, ld_l, n, rp;
    double* out;
    unsigned gx;
  };
  __device__ static void run(const Args& a, unsigned lb) {
    k_build_dfac_body(a.f0, nullptr, a.r, a.ld_l, a.n, a.rp, 1, 1.0, 0.0, a.out, lb, 0, a.gx);
  }
};
template <typename TW>
struct OpGammaW {  // k_gamma_w_partial (fp64) or k_gamma_w_blocks<4> (fp32 / bf16)
  struct Args {
    const TW* W;
    size_t kdim, rp, n, naux;
    const double* dfac;
    size_t lddf, rows;
    double* part;
    unsigned gx;
  };
  __device__ static void run(const Args& a, unsigned lb) {
    if constexpr (std::is_same<TW, double>::value)
      k_gamma_w_partial_body<TW>(a.W, nullptr, a.kdim, a.rp, a.n, a.dfac, a.lddf, a.rows, a.part, lb % a.gx, lb / a.gx, a.gx);
    else
      k_gamma_w_blocks_body<TW, 4>(a.W, nullptr, a.kdim, a.rp, a.n, a.naux, a.dfac, a.lddf, a.rows, a.part, lb % a.gx,
                                   lb / a.gx, a.gx);
  }
};
struct OpGammaReduce {
  struct Args {
    const double* part;
    int nsplit;
    size_t kdim, rp, naux;
    double* gamma;
    unsigned gx;
  };
  __device__ static void run(const Args& a, unsigned lb) {
    k_gamma_w_reduce_body(a.part, a.nsplit, a.kdim, a.rp, a.naux, a.gamma, lb, 0, a.gx);
  }
};
template <typename TB>
struct OpCoulomb {  // k_coulomb_tri (symmetric tensors) or k_coulomb, U = 4
  struct Args {
    const TB* B;
    size_t rowlen, ld, nvec, naux;
    const double* gamma;
    size_t pps;
    double* jpart;
    int tri;
    unsigned gx;
  };
  __device__ static void run(const Args& a, unsigned lb) {
    if (a.tri)
      k_coulomb_tri_body<TB, 4>(a.B, a.rowlen, a.ld, a.nvec, a.naux, a.gamma, a.pps, a.jpart, lb % a.gx, lb / a.gx, a.gx);
    else
      k_coulomb_body<TB, 4>(a.B, a.rowlen, a.naux, a.gamma, a.pps, a.jpart, lb % a.gx, lb / a.gx, a.gx);
  }
};
struct OpReduce {
  struct Args {
    const double* part;
    int nsplit;
    size_t stride, count;
    double* out;
    unsigned gx;
  };
  __device__ static void run(const Args& a, unsigned lb) {
    k_reduce_parts_body(a.part, a.nsplit, a.stride, a.count, a.out, lb, 0, a.gx);
  }
};
struct OpAsm {
  struct Args {
    RealAsm a;
    unsigned gx;
  };
  __device__ static void run(const Args& a, unsigned lb) { k_real_asm_body(a.a, lb, 0, a.gx); }
};
template <typename TI, typename TO, int BM, int BN, int BK, int WGM>
struct OpGemmLds {
  using Args = GemmArgs;
  __device__ static void run(const Args& a, unsigned lb) { gemm_nt_lds_tile<TI, TO, BM, BN, BK, WGM, 2>(a, (int)lb); }
};

// Host side of one grouped launch: the segment table, then one contiguous device copy of it.
template <typename Op>
struct Grouped {
  std::vector<typename Op::Args> args;
  std::vector<unsigned> first{0};
  const typename Op::Args* dargs = nullptr;
  const unsigned* dfirst = nullptr;
  void add(const typename Op::Args& a, size_t blocks) {
    if (!blocks) return;
    RR_REQUIRE(first.back() + blocks < (1ull << 31), RR_ERR_ARG, "batch: grouped launch too large");
    args.push_back(a);
    first.push_back(first.back() + (unsigned)blocks);
  }
  size_t bytes() const { return roundup(args.size() * sizeof(typename Op::Args), 256) + roundup(first.size() * 4, 256); }
  void upload(Context& c, char*& at) {
    if (args.empty()) return;
    const size_t ab = args.size() * sizeof(typename Op::Args);
    RR_HIP(hipMemcpyAsync(at, args.data(), ab, hipMemcpyHostToDevice, c.stream));
    dargs = reinterpret_cast<const typename Op::Args*>(at);
    at += roundup(ab, 256);
    RR_HIP(hipMemcpyAsync(at, first.data(), first.size() * 4, hipMemcpyHostToDevice, c.stream));
    dfirst = reinterpret_cast<const unsigned*>(at);
    at += roundup(first.size() * 4, 256);
  }
  void launch(hipStream_t s) const {
    if (args.empty()) return;
    hipLaunchKernelGGL(k_grouped<Op>, dim3(first.back()), dim3(256), 0, s, dargs, dfirst, (int)args.size());
  }
};

// A batch of factor-form J/K builds (rr_ri_jk_factor with D = L L^T per molecule) as grouped launches; TI is the
// operand type. The per-molecule work buffers are carved from one allocation.
template <typename TI>
struct BatchPlanT {
  Grouped<OpPack<double>> pack;
  Grouped<OpDfac> dfac;
  Grouped<OpPack<TI>> pack_t;
  // half-transform: 64 x 48 tiles (N = padded factor rank, at most a few tens here)
  // (64 x 16, 64 x 32, 128 x 16/32/48 measured slower on the 256-molecule batch: narrower tiles
  // re-read the tensor rows once per column tile)
  static constexpr int G1BM = 64, G1BN = 48;
  Grouped<OpGemmLds<TI, TI, G1BM, G1BN, 4 * Mma<TI>::KL * (std::is_same<TI, bf16_t>::value ? 2 : 1), 4>> gemm1;
  Grouped<OpGammaW<TI>> gam;
  Grouped<OpGammaReduce> gam_red;
  // rank-k update: 64 x 64 lower tiles
  Grouped<OpGemmLds<TI, double, 64, 64, 4 * Mma<TI>::KL * (std::is_same<TI, bf16_t>::value ? 2 : 1), 2>> gemm2;
  Grouped<OpCoulomb<TI>> coul;
  Grouped<OpReduce> red;
  Grouped<OpAsm> asmb;
  void* work = nullptr;
  void* tables = nullptr;
  // The Coulomb stream (reads the tensors) runs on the auxiliary stream beside the rank-k update
  // (reads W) when the context has one.
  void launch(Context& c) const {
    const hipStream_t s = c.stream;
    pack.launch(s);
    dfac.launch(s);
    pack_t.launch(s);
    gemm1.launch(s);
    gam.launch(s);
    gam_red.launch(s);
    if (c.aux) {
      RR_HIP(hipEventRecord(c.ev_fork, s));
      RR_HIP(hipStreamWaitEvent(c.aux, c.ev_fork, 0));
      coul.launch(c.aux);
      RR_HIP(hipEventRecord(c.ev_join, c.aux));
      gemm2.launch(s);
      RR_HIP(hipStreamWaitEvent(s, c.ev_join, 0));
    } else {
      gemm2.launch(s);
      coul.launch(s);
    }
    red.launch(s);
    asmb.launch(s);
  }
};

struct BatchPlan {
  int prec = P64;
  BatchPlanT<double>* p64 = nullptr;
  BatchPlanT<float>* p32 = nullptr;
  BatchPlanT<bf16_t>* p16 = nullptr;
};

namespace {
template <typename TI>
BatchPlanT<TI>* build_plan(Context& c, const std::vector<Tensor>& ts, int prec, const std::vector<size_t>& rs,
                           const double* lin, const std::vector<size_t>& loff, double* jout, double* kout,
                           const std::vector<size_t>& ooff) {
  auto* p = new BatchPlanT<TI>();
  try {
    // pass 1: work-buffer layout (bytes, 256-aligned regions)
    struct Reg {
      size_t lt, dfac, ltT, W, gpart, gamma, kpart, jpart, jr, kr;
    };
    std::vector<Reg> reg(ts.size());
    size_t at = 0;
    auto take = [&](size_t bytes) {
      const size_t o = at;
      at += roundup(std::max<size_t>(bytes, 1), 256);
      return o;
    };
    const int slab = prec == P64 ? 16 : 32;
    for (size_t m = 0; m < ts.size(); ++m) {
      const Tensor& t = ts[m];
      const size_t n = t.n, ld = t.ld, naux = t.naux, r = rs[m];
      if (r == 0 || naux == 0) continue;
      const size_t rp = roundup(r, 8), kdim = naux * rp;
      Reg& g = reg[m];
      g.lt = take(r * n * sizeof(double));
      g.dfac = take(n * rp * sizeof(double));
      g.ltT = take(rp * ld * sizeof(TI));
      g.W = take(n * kdim * sizeof(TI));
      g.gpart = take(n * (kdim / 8) * sizeof(double));  // at most n row slices
      g.gamma = take(naux * sizeof(double));
      const int ns = gemm2_splits(n, kdim, slab);
      const size_t kper = roundup((kdim + ns - 1) / ns, slab);
      const size_t nsplit = (kdim + kper - 1) / kper;
      g.kpart = take(nsplit * n * n * sizeof(double));
      g.jpart = take(32 * n * ld * sizeof(double));
      g.jr = take(n * ld * sizeof(double));
      g.kr = take(n * n * sizeof(double));
    }
    RR_HIP(hipMalloc(&p->work, std::max<size_t>(at, 256)));
    char* base = static_cast<char*>(p->work);
    // pass 2: the segments, with the grids of the single-molecule launches
    for (size_t m = 0; m < ts.size(); ++m) {
      const Tensor& t = ts[m];
      const size_t n = t.n, ld = t.ld, naux = t.naux, r = rs[m];
      double* dJ = jout + ooff[m];
      double* dK = kout + ooff[m];
      if (r == 0 || naux == 0) {  // J = K = 0
        RealAsm a;
        a.n = n, a.ld = ld, a.J = dJ, a.K = dK;
        const unsigned gx = grid_for(n * n);
        p->asmb.add({a, gx}, gx);
        continue;
      }
      const Reg& g = reg[m];
      const size_t rp = roundup(r, 8), kdim = naux * rp;
      double* lt = reinterpret_cast<double*>(base + g.lt);
      double* dfac = reinterpret_cast<double*>(base + g.dfac);
      TI* ltT = reinterpret_cast<TI*>(base + g.ltT);
      TI* W = reinterpret_cast<TI*>(base + g.W);
      double* gpart = reinterpret_cast<double*>(base + g.gpart);
      double* gamma = reinterpret_cast<double*>(base + g.gamma);
      double* kpart = reinterpret_cast<double*>(base + g.kpart);
      double* jpart = reinterpret_cast<double*>(base + g.jpart);
      double* jr = reinterpret_cast<double*>(base + g.jr);
      double* kr = reinterpret_cast<double*>(base + g.kr);
      {
        const unsigned gx = grid_for(r * n);
        p->pack.add({lin + loff[m], r, 1, r, r, n, n, 1.0, lt, gx}, gx);
      }
      {
        const unsigned gx = grid_for(n * rp);
        p->dfac.add({lt, r, n, n, rp, dfac, gx}, gx);
      }
      {
        const unsigned gx = grid_for(rp * ld);
        p->pack_t.add({lt, n, 0, rp, r, n, ld, 1.0, ltT, gx}, gx);
      }
      {  // W = Bt_P Lt^T
        GemmArgs a;
        a.A = tensor_data(t, prec), a.lda = ld, a.sAb = n * ld;
        a.B = ltT, a.ldb = ld, a.sBb = 0;
        a.C = W, a.ldcm = kdim, a.ldcn = 1, a.sCb = rp;
        // all aux rows of the molecule as one tall product (rows (P, nu) are contiguous in the
        // tensor), each row written to row nu of W's column block P
        a.M = (int)(naux * n), a.N = (int)rp, a.K = (int)roundup(n, 8), a.nbatch = 1, a.sAb = 0, a.sCb = 0;
        a.mrow = (int)n, a.sCmb = (long long)rp;
        a.tiles_m = (a.M + BatchPlanT<TI>::G1BM - 1) / BatchPlanT<TI>::G1BM;
        a.tiles_n = (a.N + BatchPlanT<TI>::G1BN - 1) / BatchPlanT<TI>::G1BN, a.ntiles = a.tiles_m * a.tiles_n;
        a.kper = a.K, a.nsplit = 1, a.xcd_remap = 0, a.zpage = c.zpage;
        p->gemm1.add(a, (size_t)a.ntiles);
      }
      int gsplit;
      {  // gamma from W (gamma_from_w)
        size_t gx, rows;
        if constexpr (std::is_same<TI, double>::value) {
          gx = ((kdim / 8) + 255) / 256;
        } else {
          gx = ((((naux + 3) / 4) * (rp / 8)) + 255) / 256;
        }
        rows = std::max<size_t>(8, (n * gx + 4095) / 4096);
        gsplit = (int)((n + rows - 1) / rows);
        p->gam.add({W, kdim, rp, n, naux, dfac, rp, rows, gpart, (unsigned)gx}, gx * (size_t)gsplit);
        const unsigned gr = gamma_reduce_grid(naux);
        p->gam_red.add({gpart, gsplit, kdim, rp, naux, gamma, gr}, gr);
      }
      int ksplit;
      {  // K partials = W W^T, lower tiles, split over k
        const int ns = gemm2_splits(n, kdim, slab);
        const size_t kper = roundup((kdim + ns - 1) / ns, slab);
        ksplit = (int)((kdim + kper - 1) / kper);
        GemmArgs a;
        a.A = W, a.lda = kdim, a.B = W, a.ldb = kdim;
        a.C = kpart, a.ldcm = n, a.ldcn = 1, a.sCz = n * n;
        a.M = (int)n, a.N = (int)n, a.K = (int)kdim, a.kper = (int)kper, a.nsplit = ksplit, a.tri = 1;
        a.tiles_m = (a.M + 63) / 64, a.tiles_n = a.tiles_m, a.ntiles = a.tiles_m * (a.tiles_m + 1) / 2;
        a.xcd_remap = 0, a.zpage = c.zpage;
        p->gemm2.add(a, (size_t)a.ntiles * ksplit);
      }
      int jsplit;
      {  // J partials (coulomb_stream)
        const bool tri = t.symmetric;
        const size_t rowlen = n * ld, vn = 16 / sizeof(TI), U = 4;
        const size_t nvec = tri ? tri_vectors(n, ld, vn) : 0;
        const size_t nx = tri ? (nvec + 256 * U - 1) / (256 * U) : (rowlen + 256 * vn * U - 1) / (256 * vn * U);
        size_t js = std::max<size_t>(1, std::min<size_t>(32, (2048 + nx - 1) / nx));
        js = std::min(js, std::max<size_t>(1, naux));
        const size_t pps = (naux + js - 1) / js;
        js = (naux + pps - 1) / pps;
        jsplit = (int)js;
        const size_t gx = std::max<size_t>(nx, 1);
        p->coul.add({static_cast<const TI*>(tensor_data(t, prec)), rowlen, ld, nvec, naux, gamma, pps, jpart, tri ? 1 : 0,
                     (unsigned)gx},
                    gx * js);
      }
      // real_assemble: reduce the split partials, then assemble J and K
      RealAsm a;
      a.n = n, a.ld = ld, a.jtri = t.symmetric ? 1 : 0;
      a.jpart = jpart, a.jsplit = jsplit, a.jstride = n * ld;
      if (jsplit > 1) {
        const unsigned gx = grid_for(n * ld);
        p->red.add({jpart, jsplit, n * ld, n * ld, jr, gx}, gx);
        a.jpart = jr, a.jsplit = 1;
      }
      a.kpart = kpart, a.ksplit = ksplit, a.kstride = n * n, a.ktile = 1;
      if (ksplit > 1) {
        const unsigned gx = grid_for(n * n);
        p->red.add({kpart, ksplit, n * n, n * n, kr, gx}, gx);
        a.kpart = kr, a.ksplit = 1;
      }
      a.kfac = -0.5, a.J = dJ, a.K = dK;
      const unsigned gx = grid_for(n * n);
      p->asmb.add({a, gx}, gx);
    }
    const size_t tb = p->pack.bytes() + p->dfac.bytes() + p->pack_t.bytes() + p->gemm1.bytes() + p->gam.bytes() +
                      p->gam_red.bytes() + p->gemm2.bytes() + p->coul.bytes() + p->red.bytes() + p->asmb.bytes();
    RR_HIP(hipMalloc(&p->tables, std::max<size_t>(tb, 256)));
    char* tat = static_cast<char*>(p->tables);
    p->pack.upload(c, tat);
    p->dfac.upload(c, tat);
    p->pack_t.upload(c, tat);
    p->gemm1.upload(c, tat);
    p->gam.upload(c, tat);
    p->gam_red.upload(c, tat);
    p->gemm2.upload(c, tat);
    p->coul.upload(c, tat);
    p->red.upload(c, tat);
    p->asmb.upload(c, tat);
    RR_HIP(hipStreamSynchronize(c.stream));  // the host tables may go once the copies are done
    return p;
  } catch (...) {
    if (p->work) (void)hipFree(p->work);
    if (p->tables) (void)hipFree(p->tables);
    delete p;
    throw;
  }
}

template <typename TI>
void free_plan(BatchPlanT<TI>* p) {
  if (!p) return;
  if (p->work) (void)hipFree(p->work);
  if (p->tables) (void)hipFree(p->tables);
  delete p;
}
}  // namespace

BatchPlan* batch_plan_create(Context& c, const std::vector<Tensor>& t, int prec, const std::vector<size_t>& r,
                             const double* lin, const std::vector<size_t>& loff, double* jout, double* kout,
                             const std::vector<size_t>& ooff) {
  auto* bp = new BatchPlan();
  bp->prec = prec;
  try {
    if (prec == P64) bp->p64 = build_plan<double>(c, t, prec, r, lin, loff, jout, kout, ooff);
    else if (prec == P32) bp->p32 = build_plan<float>(c, t, prec, r, lin, loff, jout, kout, ooff);
    else bp->p16 = build_plan<bf16_t>(c, t, prec, r, lin, loff, jout, kout, ooff);
  } catch (...) {
    delete bp;
    throw;
  }
  return bp;
}

void batch_plan_enqueue(Context& c, const BatchPlan* bp) {
  if (bp->p64) bp->p64->launch(c);
  if (bp->p32) bp->p32->launch(c);
  if (bp->p16) bp->p16->launch(c);
  RR_HIP(hipGetLastError());
}

void batch_plan_free(BatchPlan* bp) {
  if (!bp) return;
  free_plan(bp->p64);
  free_plan(bp->p32);
  free_plan(bp->p16);
  delete bp;
}

}  // namespace rr

// ================================================================================================
// C ABI
// ================================================================================================
using rr::Context;
using rr::Error;
using rr::guarded;


extern "C" {

int rr_create(int device, rr_handle* out) {
  if (!out) return RR_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count) return RR_ERR_HIP;
  rr_context* h = new rr_context();
  h->ctx.device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->ctx.stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return RR_ERR_HIP;
  }
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
      h->ctx.num_cu = cus;
  }
  if (rr::register_scratch(device) != hipSuccess) {
    (void)hipStreamDestroy(h->ctx.stream);
    delete h;
    return RR_ERR_HIP;
  }
  if (hipMalloc(&h->ctx.zpage, 256) != hipSuccess || hipMemset(h->ctx.zpage, 0, 256) != hipSuccess) {
    (void)hipStreamDestroy(h->ctx.stream);
    delete h;
    return RR_ERR_HIP;
  }
  if (rocblas_create_handle(&h->ctx.blas) != rocblas_status_success ||
      rocblas_set_stream(h->ctx.blas, h->ctx.stream) != rocblas_status_success) {
    (void)hipStreamDestroy(h->ctx.stream);
    delete h;
    return RR_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&h->ctx.aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ctx.ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ctx.ev_join, hipEventDisableTiming) != hipSuccess) {
    rr_destroy(h);
    return RR_ERR_HIP;
  }
  *out = h;
  return RR_OK;
}

int rr_destroy(rr_handle h) {
  if (!h) return RR_ERR_ARG;
  (void)hipSetDevice(h->ctx.device);
  (void)hipStreamSynchronize(h->ctx.stream);
  for (auto& t : h->ctx.slot) {
    if (t.f64) (void)hipFree(t.f64);
    if (t.f32) (void)hipFree(t.f32);
    if (t.b16) (void)hipFree(t.b16);
  }
  for (auto& r : h->ctx.rkb) {
    if (r.nat) (void)hipFree(r.nat);
    if (r.tr) (void)hipFree(r.tr);
  }
  rr::drop_plan(h->ctx.multi_plan);
  for (auto& kv : h->ctx.ws) (void)hipFree(kv.second.first);
  if (h->ctx.zpage) (void)hipFree(h->ctx.zpage);
  if (h->ctx.blas) rocblas_destroy_handle(h->ctx.blas);
  if (h->ctx.aux) {
    (void)hipStreamSynchronize(h->ctx.aux);
    (void)hipStreamDestroy(h->ctx.aux);
  }
  if (h->ctx.ev_fork) (void)hipEventDestroy(h->ctx.ev_fork);
  if (h->ctx.ev_join) (void)hipEventDestroy(h->ctx.ev_join);
  (void)hipStreamDestroy(h->ctx.stream);
  delete h;
  return RR_OK;
}

const char* rr_last_error(rr_handle h) { return h ? h->ctx.err.c_str() : "null handle"; }

int rr_set_tile_config(rr_handle h, int cfg) {
  if (!h) return RR_ERR_ARG;
  if (cfg == -1) {
    h->ctx.tile_override = -1, h->ctx.lds_override = -1;
  } else if (cfg >= 0 && cfg < rr::kNumTiles) {
    h->ctx.tile_override = cfg, h->ctx.lds_override = -2;
  } else if (cfg >= 100 && cfg < 100 + rr::kNumLds) {
    h->ctx.tile_override = -1, h->ctx.lds_override = cfg - 100;
  } else {
    h->ctx.err = "tile configuration out of range";
    return RR_ERR_ARG;
  }
  if (h->ctx.multi_plan.exec) {  // its graph holds the kernels of the previous configuration
    if (hipStreamSynchronize(h->ctx.stream) != hipSuccess) {
      h->ctx.err = "stream synchronization failed";
      return RR_ERR_HIP;
    }
    rr::drop_plan(h->ctx.multi_plan);
  }
  return RR_OK;
}

int rr_set_recovery_method(rr_handle h, int method, long long* n_projector, long long* n_eigensolver) {
  if (!h) return RR_ERR_ARG;
  if (method != -1 && method != 0 && method != 1) {
    h->ctx.err = "recovery method out of range";
    return RR_ERR_ARG;
  }
  if (method >= 0) h->ctx.recovery_method = method;
  if (n_projector) *n_projector = h->ctx.n_projector;
  if (n_eigensolver) *n_eigensolver = h->ctx.n_eigensolver;
  return RR_OK;
}

int rr_timing(rr_handle h, int enable, int reset, double* ms, long long* counts, int n) {
  if (!h) return RR_ERR_ARG;
  std::lock_guard<std::mutex> lock(rr::PhaseTimer::mu);
  rr::PhaseTimer::resolve_locked();
  for (int i = 0; i < n && i < 24; ++i) {
    if (ms) ms[i] = rr::PhaseTimer::ms[i];
    if (counts) counts[i] = rr::PhaseTimer::cnt[i];
  }
  if (reset)
    for (int i = 0; i < 24; ++i) rr::PhaseTimer::ms[i] = 0.0, rr::PhaseTimer::cnt[i] = 0;
  rr::PhaseTimer::enabled = enable != 0;
  return RR_OK;
}

int rr_tensor_release(rr_handle h, int slot) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(slot >= 0 && slot < rr::kSlots, RR_ERR_ARG, "tensor slot out of range");
    rr::Tensor& t = c.slot[slot];
    RR_HIP(hipStreamSynchronize(c.stream));
    if (t.f64) RR_HIP(hipFree(t.f64));
    if (t.f32) RR_HIP(hipFree(t.f32));
    if (t.b16) RR_HIP(hipFree(t.b16));
    t = rr::Tensor();
  });
}

static void alloc_tensor(Context& c, rr::Tensor& t, size_t naux, size_t n, unsigned copies) {
  RR_REQUIRE(copies & 7u, RR_ERR_ARG, "no tensor copy requested");
  RR_HIP(hipStreamSynchronize(c.stream));
  if (t.f64) RR_HIP(hipFree(t.f64));
  if (t.f32) RR_HIP(hipFree(t.f32));
  if (t.b16) RR_HIP(hipFree(t.b16));
  t = rr::Tensor();
  t.naux = naux, t.n = n, t.ld = rr::tensor_ld(n);
  const size_t cnt = std::max<size_t>(1, naux * n * t.ld);
  if (copies & RR_COPY_FP64) RR_HIP(hipMalloc(&t.f64, cnt * sizeof(double)));
  if (copies & RR_COPY_FP32) RR_HIP(hipMalloc(&t.f32, cnt * sizeof(float)));
  if (copies & RR_COPY_BF16) RR_HIP(hipMalloc(&t.b16, cnt * sizeof(rr::bf16_t)));
}

int rr_tensor_upload(rr_handle h, int slot, const double* eri3, size_t n_aux, size_t n, unsigned copies) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(slot >= 0 && slot < rr::kSlots, RR_ERR_ARG, "tensor slot out of range");
    rr::tensor_store(c, c.slot[slot], eri3, n_aux, n, copies);
  });
}

int rr_tensor_generate(rr_handle h, int slot, size_t n_aux, size_t n, double sigma, double decay, uint64_t seed,
                       unsigned copies) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(slot >= 0 && slot < rr::kSlots, RR_ERR_ARG, "tensor slot out of range");
    RR_REQUIRE(decay > 0.0, RR_ERR_ARG, "decay length must be positive");
    rr::Tensor& t = c.slot[slot];
    alloc_tensor(c, t, n_aux, n, copies);
    if (n_aux == 0 || n == 0) return;
    const size_t total = n_aux * n * t.ld;
    if (t.f64) hipLaunchKernelGGL(rr::k_generate<double>, dim3(rr::grid_for(total, 256, 1 << 20)), dim3(256), 0, c.stream,
                                  n_aux, n, t.ld, sigma, decay, seed, (size_t)0, n_aux, t.f64);
    if (t.f32) hipLaunchKernelGGL(rr::k_generate<float>, dim3(rr::grid_for(total, 256, 1 << 20)), dim3(256), 0, c.stream,
                                  n_aux, n, t.ld, sigma, decay, seed, (size_t)0, n_aux, t.f32);
    if (t.b16) hipLaunchKernelGGL(rr::k_generate<rr::bf16_t>, dim3(rr::grid_for(total, 256, 1 << 20)), dim3(256), 0,
                                  c.stream, n_aux, n, t.ld, sigma, decay, seed, (size_t)0, n_aux, t.b16);
    RR_HIP(hipGetLastError());
  });
}

int rr_tensor_info(rr_handle h, int slot, size_t* n_aux, size_t* n, unsigned* copies) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(slot >= 0 && slot < rr::kSlots, RR_ERR_ARG, "tensor slot out of range");
    const rr::Tensor& t = c.slot[slot];
    if (n_aux) *n_aux = t.naux;
    if (n) *n = t.n;
    if (copies) *copies = (t.f64 ? RR_COPY_FP64 : 0u) | (t.f32 ? RR_COPY_FP32 : 0u) | (t.b16 ? RR_COPY_BF16 : 0u);
  });
}

}  // extern "C"

namespace rr {
// Body of rr_ri_jk_factor (shared with its allocation-free variant, which routes the workspace
// requests to the static scratch region).
void ri_jk_factor_body(Context& c, int slot, int precision, size_t n, const double* density, const double* lfac,
                       const double* rfac, size_t r, double* j_out, double* k_out) {
  check_prec(precision);
  PrecScope prec_scope(precision);
  Tensor& t = get_slot(c, slot);
  RR_REQUIRE(t.n == n, RR_ERR_DIM, "rr_ri_jk_factor: density dimension does not match the tensor");
  double* jpart = nullptr;
  int jsplit = 0;
  const bool fused_j = j_out && !density;  // J of D = L R^T from the half-transformed factors
  RR_REQUIRE(!fused_j || (lfac && r > 0 && k_out), RR_ERR_ARG,
             "rr_ri_jk_factor: J without a density needs the factor and the exchange build");
  if (j_out && density) {
    const double* dD = in_buf(c, WS_P, density, n * n);
    jsplit = coulomb_real_prec(c, t, precision, dD, &jpart, t.symmetric);
  }
  ExchangeOut ex;
  if (k_out && r > 0) {
    RR_REQUIRE(lfac, RR_ERR_ARG, "rr_ri_jk_factor: K requested without a factor");
    const double* dL = in_buf(c, WS_CA, lfac, n * r);
    const double* dR = rfac ? in_buf(c, WS_CB, rfac, n * r) : nullptr;
    // Lt = L^T (r x n)
    double* lt = static_cast<double*>(wsbuf(c, WS_T1, 2 * r * n * sizeof(double)));
    hipLaunchKernelGGL((k_pack<double>), dim3(grid_for(r * n)), dim3(256), 0, c.stream, dL, r, 1, r, r, n, n, 1.0,
                       lt);
    if (dR)
      hipLaunchKernelGGL((k_pack<double>), dim3(grid_for(r * n)), dim3(256), 0, c.stream, dR, r, 1, r, r, n, n,
                         1.0, lt + r * n);
    RR_HIP(hipGetLastError());
    double* dfac = nullptr;
    double* gamma = nullptr;
    if (fused_j) {
      const size_t rp = roundup(r, 8);
      const int nblk = dR ? 2 : 1;
      dfac = static_cast<double*>(wsbuf(c, WS_DENS2, n * nblk * rp * sizeof(double)));
      hipLaunchKernelGGL(k_build_dfac, dim3(grid_for(n * nblk * rp)), dim3(256), 0, c.stream,
                         dR ? lt + r * n : lt, (const double*)nullptr, r, n, n, rp, nblk, 1.0, 0.0, dfac);
      RR_HIP(hipGetLastError());
      gamma = static_cast<double*>(wsbuf(c, WS_GAMMA, std::max<size_t>(1, t.naux) * sizeof(double)));
    }
    if (fused_j)
      ex = exchange_and_coulomb(c, t, precision, lt, dR ? lt + r * n : nullptr, r, n, dfac, gamma, true, t.symmetric,
                              &jpart, &jsplit);
    else
      ex = exchange_real_prec(c, t, precision, lt, dR ? lt + r * n : nullptr, r, n, dfac, gamma);
  }
  OutBuf oj = out_buf(c, WS_OUT_J, j_out, n * n), ok = out_buf(c, WS_OUT_K, k_out, n * n);
  if (k_out && r == 0) hipLaunchKernelGGL(k_zero, dim3(grid_for(n * n)), dim3(256), 0, c.stream, ok.dev, n * n);
  real_assemble(c, t, jpart, jsplit, (k_out && r > 0) ? &ex : nullptr, -0.5, nullptr, nullptr, oj.dev,
                (k_out && r > 0) ? ok.dev : nullptr, t.symmetric);
  finish_out(c, oj);
  finish_out(c, ok);
}

// Scratch region of the allocation-free entry point: one per device (each GPU the library runs on
// holds it until the process ends). Host code uses only the address obtained per device with
// hipGetSymbolAddress at context creation, never the array's own name.
constexpr size_t kNoAllocScratchBytes = size_t(160) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Records the scratch address of the current device (once per device).
hipError_t register_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[device] = p;
  return e;
}
void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}
}  // namespace rr

extern "C" {

int rr_ri_jk_factor(rr_handle h, int slot, int precision, size_t n, const double* density, const double* lfac,
                    const double* rfac, size_t r, double* j_out, double* k_out) {
  return guarded(h, [&](Context& c) { rr::ri_jk_factor_body(c, slot, precision, n, density, lfac, rfac, r, j_out, k_out); });
}

int rr_ri_jk_factor_noalloc(rr_handle h, int slot, int precision, size_t n, const double* lfac, size_t r,
                            double* j_out, double* k_out) {
  return guarded(h, [&](Context& c) {
    void* base = rr::scratch_ptr(c.device);
    RR_REQUIRE(base, RR_ERR_HIP, "rr_ri_jk_factor_noalloc: no scratch region on this device");
    RR_REQUIRE(lfac && j_out && k_out && r > 0, RR_ERR_ARG, "rr_ri_jk_factor_noalloc: needs lfac (r > 0), j_out and k_out");
    RR_REQUIRE(rr::is_device_ptr(lfac) && rr::is_device_ptr(j_out) && rr::is_device_ptr(k_out), RR_ERR_ARG,
               "rr_ri_jk_factor_noalloc: lfac, j_out and k_out must be device memory");
    rr::check_prec(precision);
    const rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == n, RR_ERR_DIM, "rr_ri_jk_factor: density dimension does not match the tensor");
    // The half-transformed factors dominate the workspace: n_aux * n * roundup(r, 8) operands.
    const size_t w_bytes = t.naux * n * rr::roundup(r, 8) * rr::prec_size(precision);
    if (w_bytes + w_bytes / 4 + (size_t(8) << 20) > rr::kNoAllocScratchBytes) {
      char msg[240];
      snprintf(msg, sizeof(msg),
               "rr_ri_jk_factor_noalloc: n_aux=%zu n=%zu r=%zu needs more than the %zu MiB scratch region "
               "(the half-transformed factors alone take %zu MiB); use rr_ri_jk_factor",
               t.naux, n, r, rr::kNoAllocScratchBytes >> 20, w_bytes >> 20);
      throw rr::Error{RR_ERR_ARG, msg};
    }
    rr::ArenaScope arena(c, base, rr::kNoAllocScratchBytes);
    rr::PhaseTimer::suppressed = true;
    struct Restore {
      ~Restore() { rr::PhaseTimer::suppressed = false; }
    } restore;
    rr::ri_jk_factor_body(c, slot, precision, n, nullptr, lfac, nullptr, r, j_out, k_out);
  });
}

int rr_nonrel_fock_ri(rr_handle h, int slot, int precision, size_t nl, const double* h_core, const double* x_large,
                      size_t nx, const double* overlap, const double* density, double* fock, double* j_out,
                      double* k_out) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == nl, RR_ERR_DIM, "nonRelFockMatrix (RI): density_matrix dimension does not match nl");
    RR_REQUIRE(density && h_core && x_large && overlap, RR_ERR_ARG, "nonRelFockMatrix (RI): null input");
    const double* dP = rr::in_buf(c, rr::WS_P, density, nl * nl);
    const double* dH = rr::in_buf(c, rr::WS_H, h_core, nl * nl);
    const double* dS = rr::in_buf(c, rr::WS_S, overlap, nl * nl);
    const double* dX = rr::in_buf(c, rr::WS_X, x_large, nl * nx);
    double* jpart = nullptr;
    const int jsplit = rr::coulomb_real_prec(c, t, precision, dP, &jpart, t.symmetric);
    // Recovery: D = P/2, C_occ = X Y_occ; c_t = sqrt(2) C_occ^T = sqrt(2) Y_occ^T X^T.
    const double* rows = nullptr;
    std::vector<double> occ;
    const size_t n_occ = rr::recover_occupied(c, false, nl, nx, dS, dP, dX, 0.5, &rows, occ, false);
    rr::ExchangeOut ex;
    if (n_occ > 0) {
      double* lt = static_cast<double*>(rr::wsbuf(c, rr::WS_COCC, n_occ * nl * sizeof(double)));
      rr::dgemm_small(c, false, 'N', 'T', n_occ, nl, nx, rows, nx, dX, nx, lt, nl, std::sqrt(2.0));  // Y^T X^T
      ex = rr::exchange_real_prec(c, t, precision, lt, nullptr, n_occ, nl);
    }
    rr::OutBuf of = rr::out_buf(c, rr::WS_OUT_F, fock, nl * nl), oj = rr::out_buf(c, rr::WS_OUT_J, j_out, nl * nl),
               ok = rr::out_buf(c, rr::WS_OUT_K, k_out, nl * nl);
    if (ok.dev && n_occ == 0) hipLaunchKernelGGL(rr::k_zero, dim3(rr::grid_for(nl * nl)), dim3(256), 0, c.stream, ok.dev, nl * nl);
    rr::real_assemble(c, t, jpart, jsplit, n_occ > 0 ? &ex : nullptr, -0.5, dH, of.dev, oj.dev, n_occ > 0 ? ok.dev : nullptr,
                      t.symmetric);
    rr::finish_out(c, of);
    rr::finish_out(c, oj);
    rr::finish_out(c, ok);
  });
}

int rr_ri_jk_multi(rr_handle h, int slot, int precision, size_t n, size_t r, const double* lfac, size_t n_dens,
                   const double* rfac, double* j_out, double* k_out, int* replayed) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == n, RR_ERR_DIM, "rr_ri_jk_multi: factor dimension does not match the tensor");
    if (replayed) *replayed = 0;
    if (n_dens == 0 || (!j_out && !k_out)) return;
    RR_REQUIRE(r > 0 && lfac && rfac, RR_ERR_ARG, "rr_ri_jk_multi: null or empty factors");
    (void)rr::tensor_data(t, precision);  // the requested copy exists
    // Inputs and outputs go through fixed staging buffers, which the captured graph reads and writes.
    double* sl = static_cast<double*>(rr::wsbuf(c, rr::WS_MU_IN_L, n * r * sizeof(double)));
    double* sr = static_cast<double*>(rr::wsbuf(c, rr::WS_MU_IN_R, n_dens * n * r * sizeof(double)));
    double* sj = j_out ? static_cast<double*>(rr::wsbuf(c, rr::WS_MU_OUT_J, n_dens * n * n * sizeof(double))) : nullptr;
    double* sk = k_out ? static_cast<double*>(rr::wsbuf(c, rr::WS_MU_OUT_K, n_dens * n * n * sizeof(double))) : nullptr;
    RR_HIP(hipMemcpyAsync(sl, lfac, n * r * sizeof(double), hipMemcpyDefault, c.stream));
    RR_HIP(hipMemcpyAsync(sr, rfac, n_dens * n * r * sizeof(double), hipMemcpyDefault, c.stream));
    const std::vector<long long> key = {slot, (long long)reinterpret_cast<uintptr_t>(rr::tensor_data(t, precision)),
                                        precision, (long long)n, (long long)r, (long long)n_dens, j_out != nullptr,
                                        k_out != nullptr, t.symmetric};
    rr::GraphPlan& plan = c.multi_plan;
    if (plan.exec && plan.key == key && plan.gen == c.ws_generation) {
      RR_HIP(hipGraphLaunch(plan.exec, c.stream));
      if (replayed) *replayed = 1;
    } else {
      rr::drop_plan(plan);
      rr::multi_jk_prec(c, t, precision, sl, sr, r, n_dens, sj, sk);  // eager: results, sizing, tuning
      RR_HIP(hipStreamSynchronize(c.stream));
      const unsigned long long gen = c.ws_generation;
      const bool timing = rr::set_phase_timing(false);
      RR_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
      try {
        rr::multi_jk_prec(c, t, precision, sl, sr, r, n_dens, sj, sk);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(c.stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        rr::set_phase_timing(timing);
        throw;
      }
      rr::set_phase_timing(timing);
      RR_HIP(hipStreamEndCapture(c.stream, &plan.graph));
      RR_REQUIRE(c.ws_generation == gen, RR_ERR_INTERNAL, "rr_ri_jk_multi: workspace reallocated during capture");
      RR_HIP(hipGraphInstantiate(&plan.exec, plan.graph, nullptr, nullptr, 0));
      plan.key = key, plan.gen = gen;
    }
    if (j_out) RR_HIP(hipMemcpyAsync(j_out, sj, n_dens * n * n * sizeof(double), hipMemcpyDefault, c.stream));
    if (k_out) RR_HIP(hipMemcpyAsync(k_out, sk, n_dens * n * n * sizeof(double), hipMemcpyDefault, c.stream));
  });
}

int rr_ri_jk_delta(rr_handle h, int slot, int precision, size_t n, size_t n_occ, const double* c_new,
                   const double* c_old, double* dj_out, double* dk_out) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == n, RR_ERR_DIM, "rr_ri_jk_delta: orbital dimension does not match the tensor");
    RR_REQUIRE(c_new && c_old, RR_ERR_ARG, "rr_ri_jk_delta: null orbitals");
    const double* dA = rr::in_buf(c, rr::WS_CA, c_new, n * n_occ);
    const double* dB = rr::in_buf(c, rr::WS_CB, c_old, n * n_occ);
    rr::OutBuf oj = rr::out_buf(c, rr::WS_OUT_J, dj_out, n * n), ok = rr::out_buf(c, rr::WS_OUT_K, dk_out, n * n);
    rr::delta_build(c, t, precision, n, n_occ, dA, dB, oj.dev, ok.dev);
    rr::finish_out(c, oj);
    rr::finish_out(c, ok);
  });
}

int rr_x2c_fock_ri(rr_handle h, int slot, int precision, size_t nl, const double* h_x2c, const double* x_large,
                   size_t nx, const double* overlap, const double* density, double* fock) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == nl, RR_ERR_DIM, "x2cFockMatrix (RI): density_matrix dimension does not match 2*nLarge");
    RR_REQUIRE(h_x2c && x_large && overlap && density, RR_ERR_ARG, "x2cFockMatrix (RI): null input");
    const size_t n = 2 * nl;
    const double* dP = rr::in_buf(c, rr::WS_P, density, 2 * n * n);
    const double* dH = rr::in_buf(c, rr::WS_H, h_x2c, 2 * n * n);
    const double* dS = rr::in_buf(c, rr::WS_S, overlap, 2 * n * n);
    const double* dX = rr::in_buf(c, rr::WS_X, x_large, 2 * n * nx);
    // Coulomb from Re(P_aa + P_bb)
    double* dl = static_cast<double*>(rr::wsbuf(c, rr::WS_DENS2, std::max<size_t>(1, nl * nl) * sizeof(double)));
    hipLaunchKernelGGL(rr::k_re_block_sum, dim3(rr::grid_for(nl * nl)), dim3(256), 0, c.stream, dP, n, 0, nl, nl, dl);
    RR_HIP(hipGetLastError());
    std::vector<const rr::Tensor*> types = {&t};
    std::vector<double*> jt = rr::coulomb_shared(c, precision, types, {dl});
    // recovery (X2C_FockMatrixRi.cpp:65-95): occupations = eigenvalues above 0.5
    const double* rows = nullptr;
    std::vector<double> occ;
    const size_t n_occ = rr::recover_occupied(c, true, n, nx, dS, dP, dX, 1.0, &rows, occ, true);
    rr::ExchangeOut ex;
    if (n_occ > 0) {
      double* ct = static_cast<double*>(rr::wsbuf(c, rr::WS_COCC, 2 * n_occ * n * sizeof(double)));
      rr::dgemm_small(c, true, 'N', 'T', n_occ, n, nx, rows, nx, dX, nx, ct, n);  // C^T = Y^T X^T
      double* docc = rr::to_device<double>(c, rr::WS_EIGW + 500, occ.data(), n_occ);
      std::vector<rr::Flavor> fl = {{&t, 0, nl}, {&t, nl, nl}};
      ex = rr::exchange_flavors_prec(c, precision, fl, ct, n, docc, n_occ, n);
    }
    rr::OutBuf of = rr::out_buf(c, rr::WS_OUT_F, fock, 2 * n * n);
    rr::cplx_assemble(c, n_occ ? &ex : nullptr, n, dH, of.dev, {{0, nl}, {nl, nl}}, {0, 0}, jt, types);
    rr::finish_out(c, of);
  });
}

int rr_ri_fock_two_electron_direct(rr_handle h, int slot_ll, int slot_ss, int precision, size_t nl, size_t ns,
                                   const double* density_ukb, const double* c_occ_ukb, size_t n_occ,
                                   const double* occupations, double* g_out) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& tl = rr::get_slot(c, slot_ll);
    rr::Tensor& ts = rr::get_slot(c, slot_ss);
    RR_REQUIRE(tl.naux == ts.naux, RR_ERR_DIM,
               "riFockTwoElectronDirect: eri3_LL and eri3_SS do not share the same aux dimension");
    RR_REQUIRE(tl.n == nl && ts.n == ns, RR_ERR_DIM,
               "riFockTwoElectronDirect: eri3_LL/eri3_SS column count does not match nl^2/ns^2");
    RR_REQUIRE(density_ukb && (c_occ_ukb || n_occ == 0), RR_ERR_ARG, "riFockTwoElectronDirect: null input");
    const size_t n = 2 * nl + 2 * ns;
    const double* dP = rr::in_buf(c, rr::WS_P, density_ukb, 2 * n * n);
    double* ct = nullptr;
    double* docc = nullptr;
    if (n_occ) {
      const double* dC = rr::in_buf(c, rr::WS_CA, c_occ_ukb, 2 * n * n_occ);
      ct = static_cast<double*>(rr::wsbuf(c, rr::WS_COCC, 2 * n_occ * n * sizeof(double)));
      hipLaunchKernelGGL(rr::k_pack_op, dim3(rr::grid_for(n_occ * n)), dim3(256), 0, c.stream, dC, n_occ, 'T', 1, n_occ, n, n,
                         ct);
      RR_HIP(hipGetLastError());
      docc = rr::to_device<double>(c, rr::WS_EIGW + 500, occupations, n_occ);
    }
    rr::OutBuf og = rr::out_buf(c, rr::WS_OUT_F, g_out, 2 * n * n);
    rr::two_electron_direct(c, precision, tl, ts, nl, ns, dP, ct, n, docc, n_occ, og.dev);
    rr::finish_out(c, og);
  });
}

int rr_rkb_fock_ri(rr_handle h, int slot_ll, int slot_ss, int precision, size_t nl, size_t ns, const double* h_rkb,
                   const double* v_total, size_t n_rkb, const double* s_full, const double* x_full, size_t nx,
                   const double* density, double* fock) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& tl = rr::get_slot(c, slot_ll);
    rr::Tensor& ts = rr::get_slot(c, slot_ss);
    RR_REQUIRE(tl.naux == ts.naux, RR_ERR_DIM,
               "riFockTwoElectronDirect: eri3_LL and eri3_SS do not share the same aux dimension");
    RR_REQUIRE(tl.n == nl && ts.n == ns, RR_ERR_DIM,
               "riFockTwoElectronDirect: eri3_LL/eri3_SS column count does not match nl^2/ns^2");
    RR_REQUIRE(h_rkb && v_total && s_full && x_full && density, RR_ERR_ARG, "rkbFockMatrix (RI): null input");
    const size_t nu = 2 * nl + 2 * ns;
    const double* dH = rr::in_buf(c, rr::WS_H, h_rkb, 2 * n_rkb * n_rkb);
    const double* dV = rr::in_buf(c, rr::WS_CB, v_total, 2 * nu * n_rkb);
    const double* dS = rr::in_buf(c, rr::WS_S, s_full, 2 * n_rkb * n_rkb);
    const double* dX = rr::in_buf(c, rr::WS_X, x_full, 2 * n_rkb * nx);
    const double* dP = rr::in_buf(c, rr::WS_P, density, 2 * n_rkb * n_rkb);
    const double* rows = nullptr;
    std::vector<double> occ;
    const size_t n_occ = rr::recover_occupied(c, true, n_rkb, nx, dS, dP, dX, 1.0, &rows, occ, true);
    double* ct_ukb = nullptr;
    double* docc = nullptr;
    if (n_occ) {
      double* ct_rkb = static_cast<double*>(rr::wsbuf(c, rr::WS_TMP8 + 300, 2 * n_occ * n_rkb * sizeof(double)));
      rr::dgemm_small(c, true, 'N', 'T', n_occ, n_rkb, nx, rows, nx, dX, nx, ct_rkb, n_rkb);      // C_rkb^T
      ct_ukb = static_cast<double*>(rr::wsbuf(c, rr::WS_COCC, 2 * n_occ * nu * sizeof(double)));
      rr::dgemm_small(c, true, 'N', 'T', n_occ, nu, n_rkb, ct_rkb, n_rkb, dV, n_rkb, ct_ukb, nu);  // (V C)^T
      docc = rr::to_device<double>(c, rr::WS_EIGW + 500, occ.data(), n_occ);
    }
    // P_ukb = V P V^H
    double* tmp = static_cast<double*>(rr::wsbuf(c, rr::WS_TMP8 + 301, 2 * n_rkb * nu * sizeof(double)));
    double* pu = static_cast<double*>(rr::wsbuf(c, rr::WS_TMP8 + 302, 2 * nu * nu * sizeof(double)));
    rr::dgemm_small(c, true, 'N', 'C', n_rkb, nu, n_rkb, dP, n_rkb, dV, n_rkb, tmp, nu);
    rr::dgemm_small(c, true, 'N', 'N', nu, nu, n_rkb, dV, n_rkb, tmp, nu, pu, nu);
    double* g = static_cast<double*>(rr::wsbuf(c, rr::WS_TMP8 + 303, 2 * nu * nu * sizeof(double)));
    rr::two_electron_direct(c, precision, tl, ts, nl, ns, pu, ct_ukb, nu, docc, n_occ, g);
    // F = h + V^H G V
    double* gv = static_cast<double*>(rr::wsbuf(c, rr::WS_TMP8 + 304, 2 * nu * n_rkb * sizeof(double)));
    rr::dgemm_small(c, true, 'N', 'N', nu, n_rkb, nu, g, nu, dV, n_rkb, gv, n_rkb);
    rr::OutBuf of = rr::out_buf(c, rr::WS_OUT_F, fock, 2 * n_rkb * n_rkb);
    hipLaunchKernelGGL(rr::k_copy_real, dim3(rr::grid_for(2 * n_rkb * n_rkb)), dim3(256), 0, c.stream, dH, of.dev,
                       2 * n_rkb * n_rkb);
    RR_HIP(hipGetLastError());
    rr::dgemm_small(c, true, 'C', 'N', n_rkb, n_rkb, nu, dV, n_rkb, gv, n_rkb, of.dev, n_rkb, 1.0, 1.0);
    rr::finish_out(c, of);
  });
}

int rr_build_ri_mo_eri(rr_handle h, int slot_ll, int slot_ss, int precision, int out_precision, size_t nl, size_t ns,
                       const double* c_ukb, size_t n_total, void* out) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    RR_REQUIRE(out_precision == RR_FP64 || out_precision == RR_FP32, RR_ERR_ARG, "buildRiMoEri: output precision");
    RR_REQUIRE(!(precision == RR_FP64 && out_precision == RR_FP32), RR_ERR_ARG,
               "buildRiMoEri: FP64 arithmetic writes an FP64 tensor");
    rr::Tensor& tl = rr::get_slot(c, slot_ll);
    RR_REQUIRE(tl.n == nl, RR_ERR_DIM, "buildRiMoEri: c_ukb row count does not match 2*nLarge+2*nSmall");
    const rr::Tensor* ts = nullptr;
    if (ns > 0) {
      ts = &rr::get_slot(c, slot_ss);
      RR_REQUIRE(ts->n == ns && ts->naux == tl.naux, RR_ERR_DIM,
                 "buildRiMoEri: c_ukb row count does not match 2*nLarge+2*nSmall");
    }
    RR_REQUIRE(c_ukb && out, RR_ERR_ARG, "buildRiMoEri: null input");
    const size_t nu = 2 * nl + 2 * ns;
    const double* dC = rr::in_buf(c, rr::WS_CA, c_ukb, 2 * nu * n_total);
    std::vector<rr::Flavor> fl = {{&tl, 0, nl}, {&tl, nl, nl}};
    if (ns > 0) fl.push_back({ts, 2 * nl, ns}), fl.push_back({ts, 2 * nl + ns, ns});
    const size_t count = 2 * tl.naux * n_total * n_total;
    const size_t esz = out_precision == RR_FP64 ? 8 : 4;
    const bool dev = rr::is_device_ptr(out);
    if (dev) RR_VALIDATE_EXTENT(out, count * esz, "buildRiMoEri output");
    void* dout = dev ? out : rr::wsbuf(c, rr::WS_MO_OUT, std::max<size_t>(1, count) * esz);
    if (precision == RR_FP64) rr::mo_cplx_slot<double, double>(c, fl, precision, dC, n_total, nu, (double*)dout);
    else if (precision == RR_FP32) rr::mo_cplx_slot<float, float>(c, fl, precision, dC, n_total, nu, (float*)dout);
    else if (out_precision == RR_FP32) rr::mo_cplx_slot<rr::bf16_t, float>(c, fl, precision, dC, n_total, nu, (float*)dout);
    else rr::mo_cplx_slot<rr::bf16_t, double>(c, fl, precision, dC, n_total, nu, (double*)dout);
    if (!dev && count) RR_HIP(hipMemcpyAsync(out, dout, count * esz, hipMemcpyDeviceToHost, c.stream));
  });
}

int rr_ri_mo_eri_rotated(rr_handle h, int precision, size_t n_total, size_t n_aux, const double* b, const double* u,
                         double* out) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    RR_REQUIRE(b && u && out, RR_ERR_ARG, "RiMoEri::rotated: null input");
    const size_t cnt = 2 * n_aux * n_total * n_total;
    const double* db = rr::in_buf(c, rr::WS_P, b, cnt);
    const double* du = rr::in_buf(c, rr::WS_X, u, 2 * n_total * n_total);
    rr::OutBuf oo = rr::out_buf(c, rr::WS_OUT_F, out, cnt);
    if (precision == RR_FP64) rr::rotate_cplx<double>(c, precision, n_total, n_aux, db, du, oo.dev);
    else if (precision == RR_FP32) rr::rotate_cplx<float>(c, precision, n_total, n_aux, db, du, oo.dev);
    else rr::rotate_cplx<rr::bf16_t>(c, precision, n_total, n_aux, db, du, oo.dev);
    rr::finish_out(c, oo);
  });
}

int rr_ri_mo_eri_positive_block(rr_handle h, size_t n_total, size_t n_aux, const double* b, size_t off, double* out) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(off <= n_total, RR_ERR_ARG, "RiMoEri::positiveBlock: offset beyond n_total");
    RR_REQUIRE(b && out, RR_ERR_ARG, "RiMoEri::positiveBlock: null input");
    const size_t m = n_total - off;
    const double* db = rr::in_buf(c, rr::WS_P, b, 2 * n_aux * n_total * n_total);
    rr::OutBuf oo = rr::out_buf(c, rr::WS_OUT_F, out, 2 * n_aux * m * m);
    hipLaunchKernelGGL(rr::k_positive_block, dim3(rr::grid_for(n_aux * m * m)), dim3(256), 0, c.stream, db, n_total, n_aux, off,
                       oo.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, oo);
  });
}

int rr_build_ri_nonrel_spin_mo_eri(rr_handle h, int slot, int precision, size_t n, const double* c_spatial, size_t n_mo,
                                   double* out) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == n, RR_ERR_DIM, "buildRiNonRelSpinMoEri: eri3_L column count does not match n_spatial*n_spatial");
    RR_REQUIRE(c_spatial && out, RR_ERR_ARG, "buildRiNonRelSpinMoEri: null input");
    const double* dC = rr::in_buf(c, rr::WS_CA, c_spatial, n * n_mo);
    rr::OutBuf oo = rr::out_buf(c, rr::WS_MO_OUT, out, t.naux * n_mo * n_mo);
    if (precision == RR_FP64) rr::mo_real_slot<double, double>(c, t, precision, dC, n_mo, oo.dev);
    else if (precision == RR_FP32) rr::mo_real_slot<float, double>(c, t, precision, dC, n_mo, oo.dev);
    else rr::mo_real_slot<rr::bf16_t, double>(c, t, precision, dC, n_mo, oo.dev);
    rr::finish_out(c, oo);
  });
}

int rr_ri_nonrel_spin_mo_eri_rotated(rr_handle h, int precision, size_t n_spatial, size_t n_aux, const double* b,
                                     const double* u, double* out) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    RR_REQUIRE(b && u && out, RR_ERR_ARG, "RiNonRelSpinMoEri::rotated: null input");
    const size_t n = n_spatial, n2 = 2 * n;
    std::vector<double> uh(n2 * n2);
    RR_HIP(hipMemcpyAsync(uh.data(), u, n2 * n2 * sizeof(double), hipMemcpyDefault, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
    double max_off = 0.0, max_mis = 0.0;
    std::vector<double> us(n * n);
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) {
        max_off = std::max({max_off, std::fabs(uh[i * n2 + n + j]), std::fabs(uh[(n + i) * n2 + j])});
        max_mis = std::max(max_mis, std::fabs(uh[(n + i) * n2 + n + j] - uh[i * n2 + j]));
        us[i * n + j] = 0.5 * (uh[i * n2 + j] + uh[(n + i) * n2 + n + j]);
      }
    if (max_off > 1e-3 || max_mis > 1e-3) {
      char buf[256];
      snprintf(buf, sizeof buf,
               "RiNonRelSpinMoEri::rotated: rotation is not spin-restricted (alpha and beta rotations differ by %f, "
               "cross-spin block max %f)",
               max_mis, max_off);
      throw rr::Error{RR_ERR_ARG, buf};
    }
    const double* db = rr::in_buf(c, rr::WS_P, b, n_aux * n * n);
    const double* du = rr::to_device<double>(c, rr::WS_X, us.data(), n * n);
    rr::OutBuf oo = rr::out_buf(c, rr::WS_OUT_F, out, n_aux * n * n);
    if (precision == RR_FP64) rr::rotate_real<double>(c, precision, n, n_aux, db, du, oo.dev);
    else if (precision == RR_FP32) rr::rotate_real<float>(c, precision, n, n_aux, db, du, oo.dev);
    else rr::rotate_real<rr::bf16_t>(c, precision, n, n_aux, db, du, oo.dev);
    rr::finish_out(c, oo);
  });
}

int rr_ri_nonrel_spin_embedded_b(rr_handle h, size_t n_spatial, size_t n_aux, const double* b, double* out) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b && out, RR_ERR_ARG, "RiNonRelSpinMoEri::embeddedB: null input");
    const size_t n2 = 2 * n_spatial;
    const double* db = rr::in_buf(c, rr::WS_P, b, n_aux * n_spatial * n_spatial);
    rr::OutBuf oo = rr::out_buf(c, rr::WS_OUT_F, out, n_aux * n2 * n2);
    hipLaunchKernelGGL(rr::k_embedded_b, dim3(rr::grid_for(n_aux * n2 * n2)), dim3(256), 0, c.stream, db, n_spatial, n_aux,
                       oo.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, oo);
  });
}

int rr_nonrel_fock_cholesky(rr_handle h, int slot, int precision, size_t n, const double* h_core, const double* p,
                            double* fock) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == n, RR_ERR_DIM, "nonRelFockMatrix: AO Cholesky dimension does not match");
    RR_REQUIRE(h_core && p, RR_ERR_ARG, "nonRelFockMatrix: density dimensions inconsistent");
    const double* dP = rr::in_buf(c, rr::WS_P, p, n * n);
    const double* dH = rr::in_buf(c, rr::WS_H, h_core, n * n);
    // J: t_L = sum_qs B(q,s) P(s,q) = <Bt_L, P>, i.e. the Coulomb pass with D = P^T
    double* pt = static_cast<double*>(rr::wsbuf(c, rr::WS_DENS2, n * n * sizeof(double)));
    hipLaunchKernelGGL((rr::k_pack<double>), dim3(rr::grid_for(n * n)), dim3(256), 0, c.stream, dP, n, 1, n, n, n, n, 1.0, pt);
    double* jpart = nullptr;
    const int js = rr::coulomb_real_prec(c, t, precision, pt, &jpart);
    double* k = static_cast<double*>(rr::wsbuf(c, rr::WS_OUT_K, n * n * sizeof(double)));
    rr::cholesky_exchange_prec(c, t, precision, dP, n, k, n, false);
    rr::OutBuf of = rr::out_buf(c, rr::WS_OUT_F, fock, n * n);
    hipLaunchKernelGGL(rr::k_fock_real_k, dim3(rr::grid_for(n * n)), dim3(256), 0, c.stream, n, t.ld, jpart, js, n * t.ld,
                       k, dH, of.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, of);
  });
}

int rr_x2c_fock_cholesky(rr_handle h, int slot, int precision, size_t n, const double* h_x2c, const double* p,
                         double* fock) {
  return guarded(h, [&](Context& c) {
    rr::check_prec(precision);
    rr::PrecScope prec_scope(precision);
    rr::Tensor& t = rr::get_slot(c, slot);
    RR_REQUIRE(t.n == n, RR_ERR_DIM, "x2cFockMatrix: AO Cholesky dimension does not match");
    RR_REQUIRE(h_x2c && p, RR_ERR_ARG, "x2cFockMatrix: density dimensions inconsistent");
    const size_t n2 = 2 * n;
    const double* dP = rr::in_buf(c, rr::WS_P, p, 2 * n2 * n2);
    const double* dH = rr::in_buf(c, rr::WS_H, h_x2c, 2 * n2 * n2);
    // J: t_L = sum_s sum_qr B(q,r) Re P_ss(r,q)  -> Coulomb with D = (Re(P_aa + P_bb))^T
    double* dl = static_cast<double*>(rr::wsbuf(c, rr::WS_DENS2, 2 * n * n * sizeof(double)));
    hipLaunchKernelGGL(rr::k_re_block_sum, dim3(rr::grid_for(n * n)), dim3(256), 0, c.stream, dP, n2, 0, n, n, dl);
    hipLaunchKernelGGL((rr::k_pack<double>), dim3(rr::grid_for(n * n)), dim3(256), 0, c.stream, dl, n, 1, n, n, n, n, 1.0,
                       dl + n * n);
    RR_HIP(hipGetLastError());
    double* jpart = nullptr;
    const int js = rr::coulomb_real_prec(c, t, precision, dl + n * n, &jpart);
    double* jt = static_cast<double*>(rr::wsbuf(c, rr::WS_OUT_J, n * t.ld * sizeof(double)));
    hipLaunchKernelGGL(rr::k_reduce_parts, dim3(rr::grid_for(n * t.ld)), dim3(256), 0, c.stream, jpart, js, n * t.ld,
                       n * t.ld, jt);
    double* ksu = static_cast<double*>(rr::wsbuf(c, rr::WS_OUT_K, 8 * n * n * sizeof(double)));
    double* pop = static_cast<double*>(rr::wsbuf(c, rr::WS_TMP8 + 310, 2 * n * n * sizeof(double)));
    for (int s = 0; s < 2; ++s)
      for (int u = 0; u < 2; ++u) {
        hipLaunchKernelGGL(rr::k_split_block, dim3(rr::grid_for(2 * n * n)), dim3(256), 0, c.stream, dP, n, (size_t)s,
                           (size_t)u, pop);
        RR_HIP(hipGetLastError());
        rr::cholesky_exchange_prec(c, t, precision, pop, 2 * n, ksu + (s * 2 + u) * 2 * n * n, 2 * n, true);
      }
    rr::OutBuf of = rr::out_buf(c, rr::WS_OUT_F, fock, 2 * n2 * n2);
    hipLaunchKernelGGL(rr::k_fock_x2c_chol, dim3(rr::grid_for(n2 * n2)), dim3(256), 0, c.stream, n, t.ld, jt, ksu, dH, of.dev);
    RR_HIP(hipGetLastError());
    rr::finish_out(c, of);
  });
}

int rr_cholesky_to_mo_spatial(rr_handle h, int slot, int precision, size_t n, const double* cm, size_t n_mo,
                              double* out) {
  return rr_build_ri_nonrel_spin_mo_eri(h, slot, precision, n, cm, n_mo, out);
}

int rr_cholesky_to_mo_spinor(rr_handle h, int slot, int precision, size_t n, const double* cm, size_t n_mo,
                             double* out) {
  // W_L = (sum_s C_s^H B_L C_s)^T = conj(sum_s C_s^H B_L C_s) for symmetric B_L: the X2C two-block
  // MO transform followed by conjugation.
  int rc = rr_build_ri_mo_eri(h, slot, slot, precision, RR_FP64, n, 0, cm, n_mo, out);
  if (rc != RR_OK) return rc;
  return guarded(h, [&](Context& c) {
    const rr::Tensor& t = rr::get_slot(c, slot);
    const size_t cnt = t.naux * n_mo * n_mo;
    if (rr::is_device_ptr(out)) {
      hipLaunchKernelGGL(rr::k_conj_inplace, dim3(rr::grid_for(cnt)), dim3(256), 0, c.stream, out, cnt);
      RR_HIP(hipGetLastError());
    } else {
      for (size_t i = 0; i < cnt; ++i) out[2 * i + 1] = -out[2 * i + 1];
    }
  });
}

}  // extern "C"
