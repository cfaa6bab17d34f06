// Device-side primitives shared by every gfx950 kernel of the RI J/K library: vector types, the
// three CDNA4 matrix-core instructions the GEMM engine is built on, and bf16 conversion.
#pragma once

#include <hip/hip_runtime.h>

#include "rr_validate.hpp"

#include <cstdint>

namespace rr {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// bf16 storage type (raw bits); conversion rounds to nearest even.
struct bf16_t {
  uint16_t bits;
};

__host__ __device__ inline float bf16_to_f32(bf16_t v) {
  union {
    uint32_t u;
    float f;
  } x;
  x.u = static_cast<uint32_t>(v.bits) << 16;
  return x.f;
}

__host__ __device__ inline bf16_t f32_to_bf16(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  // one v_cvt_pk_bf16_f32 (round to nearest even); the branchy software form below cost about
  // twenty instructions per value in the GEMM epilogues
  bf16_t r;
  r.bits = __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
  return r;
#else
  union {
    uint32_t u;
    float f;
  } x;
  x.f = f;
  uint32_t u = x.u;
  if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: truncate, keep nan quiet
    bf16_t r;
    r.bits = static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    return r;
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  bf16_t r;
  r.bits = static_cast<uint16_t>(u >> 16);
  return r;
#endif
}

template <typename T>
struct Conv;
template <>
struct Conv<double> {
  __device__ static double to(double v) { return v; }
  __device__ static double from(double v) { return v; }
};
template <>
struct Conv<float> {
  __device__ static float to(double v) { return static_cast<float>(v); }
  __device__ static double from(float v) { return static_cast<double>(v); }
};
template <>
struct Conv<bf16_t> {
  __device__ static bf16_t to(double v) { return f32_to_bf16(static_cast<float>(v)); }
  __device__ static double from(bf16_t v) { return static_cast<double>(bf16_to_f32(v)); }
};

// ------------------------------------------------------------------------------------------------
// Matrix-core traits. For every instruction used here, a 64-lane wave computes a 16x16 output tile;
// lane l owns four outputs of column (l & 15), in rows out_row(l >> 4, e), e = 0..3. Operand
// fragments: lane l supplies row (l & 15) of A (and of B^T) at the k-slot of its lane group
// g = l >> 4.
//
// The engine gives every lane a contiguous run of KL k-elements per "super-step" (32 bytes for
// fp64/fp32, 16 bytes for bf16). Lane group g owns k-run [g*KL, (g+1)*KL) of the super-step's
// 4*KL-wide k-slab, so each instruction s of the super-step contracts the four k's
// {g*KL + s : g = 0..3} (fp64/fp32) or the four 8-wide blocks {g*KL + [0,8)} (bf16). The sum over
// the slab is complete and each k is used once: the permutation of k is invisible to the result.
// ------------------------------------------------------------------------------------------------
template <typename TI>
struct Mma;

template <>
struct Mma<double> {
  using acc_t = d4;
  using acc_scalar = double;
  // Measured on gfx950 (one-hot operand probe): the f64 instruction returns D[i][j] in lane
  // j + 16*(i%4), register i/4 -- rows interleaved across lane groups, unlike the f32/bf16 forms.
  __device__ static int out_row(int grp, int e) { return 4 * e + grp; }
  static constexpr int KL = 4;       // elements per lane per super-step
  static constexpr int STEPS = 4;    // instructions per super-step
  struct frag {
    double v[4];
  };
  __device__ static void load(frag& f, const double* p) {
    const d2* q = reinterpret_cast<const d2*>(p);
    d2 a = q[0], b = q[1];
    f.v[0] = a.x, f.v[1] = a.y, f.v[2] = b.x, f.v[3] = b.y;
  }
  __device__ static void zero(frag& f) { f.v[0] = f.v[1] = f.v[2] = f.v[3] = 0.0; }
  __device__ static void mma(acc_t& c, const frag& a, const frag& b) {
#pragma unroll
    for (int s = 0; s < 4; ++s) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v[s], b.v[s], c, 0, 0, 0);
  }
  // All tiles of a wave, step-major so consecutive instructions are independent.
  template <int WM, int WN>
  __device__ static void mma_tiles(acc_t (&c)[WM][WN], const frag (&a)[WM], const frag (&b)[WN]) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) c[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].v[s], b[j].v[s], c[i][j], 0, 0, 0);
  }
};

template <>
struct Mma<float> {
  using acc_t = f4;
  using acc_scalar = float;
  __device__ static int out_row(int grp, int e) { return 4 * grp + e; }
  static constexpr int KL = 8;
  static constexpr int STEPS = 8;
  struct frag {
    float v[8];
  };
  __device__ static void load(frag& f, const float* p) {
    const f4* q = reinterpret_cast<const f4*>(p);
    f4 a = q[0], b = q[1];
    f.v[0] = a.x, f.v[1] = a.y, f.v[2] = a.z, f.v[3] = a.w;
    f.v[4] = b.x, f.v[5] = b.y, f.v[6] = b.z, f.v[7] = b.w;
  }
  __device__ static void zero(frag& f) {
#pragma unroll
    for (int i = 0; i < 8; ++i) f.v[i] = 0.f;
  }
  __device__ static void mma(acc_t& c, const frag& a, const frag& b) {
#pragma unroll
    for (int s = 0; s < 8; ++s) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[s], b.v[s], c, 0, 0, 0);
  }
  template <int WM, int WN>
  __device__ static void mma_tiles(acc_t (&c)[WM][WN], const frag (&a)[WM], const frag (&b)[WN]) {
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) c[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].v[s], b[j].v[s], c[i][j], 0, 0, 0);
  }
};

template <>
struct Mma<bf16_t> {
  using acc_t = f4;
  using acc_scalar = float;
  __device__ static int out_row(int grp, int e) { return 4 * grp + e; }
  static constexpr int KL = 8;
  static constexpr int STEPS = 1;
  struct frag {
    bf8 v;
  };
  __device__ static void load(frag& f, const bf16_t* p) { f.v = *reinterpret_cast<const bf8*>(p); }
  __device__ static void zero(frag& f) {
    f.v = bf8{};
  }
  __device__ static void mma(acc_t& c, const frag& a, const frag& b) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
  }
  template <int WM, int WN>
  __device__ static void mma_tiles(acc_t (&c)[WM][WN], const frag (&a)[WM], const frag (&b)[WN]) {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j) c[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i].v, b[j].v, c[i][j], 0, 0, 0);
  }
};

}  // namespace rr
