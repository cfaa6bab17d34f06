// The MFMA GEMM engine: C(m,n) = alpha * sum_k opA(A(m,k)) * opB(B(n,k)) [+ beta * C(m,n)].
//
// Every contraction of the RI hot path is written in this one "NT" form -- both operands stored
// row-major with k contiguous -- by choosing the device layouts of the 3-index tensor and of its
// half-transformed intermediate accordingly (rr_lib.hip). Operand fragments are loaded straight from
// global memory (L2/MALL-resident panels) into registers in 32-byte (fp64/fp32) or 16-byte (bf16)
// runs, so the kernel needs no LDS and no barriers; each 256-thread workgroup is 2x2 waves and each
// wave owns a (16*WM) x (16*WN) tile of 16x16 matrix-core accumulators.
//
// Features: batched operands (batch strides), split-K into separate partial outputs, lower-triangle
// tile scheduling for Hermitian rank-k updates, complex operands (interleaved re/im) handled as real
// GEMMs with the sign/swap of the B fragment chosen by the conjugation mode, and an XCD-aware
// workgroup order so that the tiles of one k-split share one XCD's L2.
#pragma once

#include "rr_device.hpp"

namespace rr {

struct GemmArgs {
  const void* A = nullptr;
  long long lda = 0, sAb = 0;  // logical elements (a complex value counts once)
  const void* B = nullptr;
  long long ldb = 0, sBb = 0;
  void* C = nullptr;
  long long ldcm = 0, ldcn = 1, sCb = 0, sCz = 0;
  int M = 0, N = 0, K = 0;
  int kper = 0;          // logical k per split (multiple of the k-slab)
  int tiles_m = 0, tiles_n = 0, ntiles = 0;
  int tri = 0;           // 1: only tiles with tile_m >= tile_n
  int nbatch = 1, nsplit = 1;
  int xcd_remap = 1;
  double alpha = 1.0, beta = 0.0;
  int conj = 0;          // bit0: conjugate A, bit1: conjugate B (complex only)
  // Optional second output: rows m >= mcut (if mcut > 0) or columns n >= ncut (if ncut > 0) are
  // written to C2 at (m - mcut, n) or (m, n - ncut) with the same strides.
  void* C2 = nullptr;
  int mcut = 0, ncut = 0;
  // Column blocks: with nblk > 0, column n goes to block n / nblk (at sCnb logical elements from
  // block 0) as column n % nblk -- several outputs of one launch that share A.
  int nblk = 0;
  long long sCnb = 0;
  // k-batched reduction: the contraction runs over kbatch blocks of K each; block b of A (B) starts
  // sAk (sBk) logical elements after block 0. kper then counts flat k over kbatch * K.
  int kbatch = 0;
  long long sAk = 0, sBk = 0;
  // Row blocks of the output (gemm_nt_lds_kernel only): with mrow > 0, row m is written as row
  // m % mrow of block m / mrow, blocks sCmb logical elements apart -- a stack of products that share
  // the B operand computed as one tall product.
  int mrow = 0;
  long long sCmb = 0;
  // 16+ bytes of zeros in device memory: the source of out-of-range chunks in the direct-to-LDS kernel
  const void* zpage = nullptr;
};


template <typename TO>
__device__ inline void store_out(TO* p, double v);
template <>
__device__ inline void store_out<double>(double* p, double v) {
  *p = v;
}
template <>
__device__ inline void store_out<float>(float* p, double v) {
  *p = static_cast<float>(v);
}
template <>
__device__ inline void store_out<bf16_t>(bf16_t* p, double v) {
  *p = f32_to_bf16(static_cast<float>(v));
}

__device__ inline int remap_xcd(int w, int total) {
  // Workgroup w runs on XCD (w % 8). Give XCD x a contiguous block of logical work items.
  const int per = total >> 3, rem = total & 7, x = w & 7;
  return x * per + (x < rem ? x : rem) + (w >> 3);
}

// Complex fragments hold interleaved (re, im) pairs.
__device__ inline void cplx_conj(Mma<double>::frag& f) {
  f.v[1] = -f.v[1];
  f.v[3] = -f.v[3];
}
__device__ inline void cplx_conj(Mma<float>::frag& f) {
#pragma unroll
  for (int p = 1; p < 8; p += 2) f.v[p] = -f.v[p];
}
typedef unsigned short us8 __attribute__((ext_vector_type(8)));
__device__ inline void cplx_conj(Mma<bf16_t>::frag& f) {
  us8 u = __builtin_bit_cast(us8, f.v);
  u ^= us8{0, 0x8000, 0, 0x8000, 0, 0x8000, 0, 0x8000};
  f.v = __builtin_bit_cast(bf8, u);
}
template <typename F>
__device__ inline void cplx_split_pairs(const F& b, F& fr, F& fi, bool conj_b, int npairs) {
#pragma unroll
  for (int p = 0; p < npairs; ++p) {
    const auto br = b.v[2 * p], bi = b.v[2 * p + 1];
    fr.v[2 * p] = br;
    fr.v[2 * p + 1] = conj_b ? bi : -bi;
    fi.v[2 * p] = conj_b ? -bi : bi;
    fi.v[2 * p + 1] = br;
  }
}
__device__ inline void cplx_split(const Mma<double>::frag& b, Mma<double>::frag& fr, Mma<double>::frag& fi, bool cb) {
  cplx_split_pairs(b, fr, fi, cb, 2);
}
__device__ inline void cplx_split(const Mma<float>::frag& b, Mma<float>::frag& fr, Mma<float>::frag& fi, bool cb) {
  cplx_split_pairs(b, fr, fi, cb, 4);
}
__device__ inline void cplx_split(const Mma<bf16_t>::frag& b, Mma<bf16_t>::frag& fr, Mma<bf16_t>::frag& fi, bool cb) {
  const us8 u = __builtin_bit_cast(us8, b.v);
  const unsigned short S = 0x8000;
  us8 r, i;
  // fr = (br, cb ? bi : -bi), fi = (cb ? -bi : bi, br)
  r = us8{u[0], (unsigned short)(u[1] ^ (cb ? 0 : S)), u[2], (unsigned short)(u[3] ^ (cb ? 0 : S)),
          u[4], (unsigned short)(u[5] ^ (cb ? 0 : S)), u[6], (unsigned short)(u[7] ^ (cb ? 0 : S))};
  i = us8{(unsigned short)(u[1] ^ (cb ? S : 0)), u[0], (unsigned short)(u[3] ^ (cb ? S : 0)), u[2],
          (unsigned short)(u[5] ^ (cb ? S : 0)), u[4], (unsigned short)(u[7] ^ (cb ? S : 0)), u[6]};
  fr.v = __builtin_bit_cast(bf8, r);
  fi.v = __builtin_bit_cast(bf8, i);
}

// One workgroup's tile of gemm_nt_kernel; w = linear workgroup index (tile, batch, split).
template <typename TI, typename TO, int WM, int WN, bool CPLX>
__device__ __forceinline__ void gemm_nt_tile(const GemmArgs& g, int w) {
  using MM = Mma<TI>;
  using acc_t = typename MM::acc_t;
  using frag = typename MM::frag;
  constexpr int KL = MM::KL;
  constexpr int SLAB = 4 * KL;  // real elements of k per super-step
  constexpr int CW = CPLX ? 2 : 1;
  constexpr int BM = 32 * WM, BN = 32 * WN;

  const int tile = w % g.ntiles;
  const int rest = w / g.ntiles;
  const int batch = rest % g.nbatch;
  const int split = rest / g.nbatch;

  int tm, tn;
  if (g.tri) {
    int t = tile;
    int a = static_cast<int>((sqrtf(8.0f * static_cast<float>(t) + 1.0f) - 1.0f) * 0.5f);
    while ((a + 1) * (a + 2) / 2 <= t) ++a;
    while (a * (a + 1) / 2 > t) --a;
    tm = a;
    tn = t - a * (a + 1) / 2;
  } else {
    tm = tile % g.tiles_m;
    tn = tile / g.tiles_m;
  }

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, grp = lane >> 4;
  const int m0 = tm * BM + (wave & 1) * (16 * WM);
  const int n0 = tn * BN + (wave >> 1) * (16 * WN);

  const long long kbeg = static_cast<long long>(split) * g.kper * CW;
  long long kend = kbeg + static_cast<long long>(g.kper) * CW;
  const long long kblk = static_cast<long long>(g.K) * CW;  // real elements per k-block
  const long long ktot = kblk * (g.kbatch > 0 ? g.kbatch : 1);
  if (kend > ktot) kend = ktot;

  const TI* Ab = static_cast<const TI*>(g.A) + static_cast<long long>(batch) * g.sAb * CW;
  const TI* Bb = static_cast<const TI*>(g.B) + static_cast<long long>(batch) * g.sBb * CW;
  const TI* arow[WM];
  const TI* brow[WN];
#pragma unroll
  for (int i = 0; i < WM; ++i) {
    const int m = m0 + 16 * i + r;
    arow[i] = (m < g.M) ? Ab + static_cast<long long>(m) * g.lda * CW : nullptr;
  }
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const int n = n0 + 16 * j + r;
    brow[j] = (n < g.N) ? Bb + static_cast<long long>(n) * g.ldb * CW : nullptr;
  }

  acc_t acc[WM][WN];
  acc_t acci[CPLX ? WM : 1][CPLX ? WN : 1];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = acc_t{};
  if constexpr (CPLX) {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j) acci[i][j] = acc_t{};
  }

  const bool ca = (g.conj & 1) != 0, cb = (g.conj & 2) != 0;
  for (long long k = kbeg; k < kend; k += SLAB) {
    const long long kk = k + grp * KL;
    const bool kin = kk < kend;
    long long ka = kk, kb = kk;
    if (g.kbatch > 0) {
      const long long blk = kk / kblk, kl = kk - blk * kblk;
      ka = blk * g.sAk * CW + kl;
      kb = blk * g.sBk * CW + kl;
    }
    frag fa[WM], fb[WN];
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      if (kin && arow[i])
        MM::load(fa[i], arow[i] + ka);
      else
        MM::zero(fa[i]);
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      if (kin && brow[j])
        MM::load(fb[j], brow[j] + kb);
      else
        MM::zero(fb[j]);
    }
    if constexpr (!CPLX) {
      MM::mma_tiles(acc, fa, fb);
    } else {
      // Interleaved (re, im) pairs; conj(A) is applied to the A fragment itself, then
      // Re C = a . bR and Im C = a . bI with, per pair (br, bi):
      //   B plain: bR = (br, -bi), bI = ( bi, br)
      //   conj(B): bR = (br,  bi), bI = (-bi, br)
      if (ca) {
#pragma unroll
        for (int i = 0; i < WM; ++i) cplx_conj(fa[i]);
      }
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        frag fr, fi;
        cplx_split(fb[j], fr, fi, cb);
#pragma unroll
        for (int i = 0; i < WM; ++i) {
          MM::mma(acc[i][j], fa[i], fr);
          MM::mma(acci[i][j], fa[i], fi);
        }
      }
    }
  }

  // Epilogue: lane owns rows out_row(grp, e) of column r in each 16x16 tile.
  const long long coff = (static_cast<long long>(batch) * g.sCb + static_cast<long long>(split) * g.sCz) * CW;
  TO* Cb = static_cast<TO*>(g.C) + coff;
  TO* Cb2 = g.C2 ? static_cast<TO*>(g.C2) + coff : nullptr;
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int n = n0 + 16 * j + r;
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = m0 + 16 * i + MM::out_row(grp, e);
        if (m >= g.M) continue;
        int mm = m, nn = n;
        TO* base = Cb;
        if (g.mcut > 0 && mm >= g.mcut) mm -= g.mcut, base = Cb2;
        if (g.ncut > 0 && nn >= g.ncut) nn -= g.ncut, base = Cb2;
        if (g.nblk > 0) {
          const int b = nn / g.nblk;
          nn -= b * g.nblk;
          base += static_cast<long long>(b) * g.sCnb * CW;
        }
        TO* p = base + (static_cast<long long>(mm) * g.ldcm + static_cast<long long>(nn) * g.ldcn) * CW;
        double v = g.alpha * static_cast<double>(acc[i][j][e]);
        if (g.beta != 0.0) v += g.beta * Conv<TO>::from(p[0]);
        store_out<TO>(p, v);
        if constexpr (CPLX) {
          double vi = g.alpha * static_cast<double>(acci[i][j][e]);
          if (g.beta != 0.0) vi += g.beta * Conv<TO>::from(p[1]);
          store_out<TO>(p + 1, vi);
        }
      }
    }
}

template <typename TI, typename TO, int WM, int WN, bool CPLX>
__global__ __launch_bounds__(256) void gemm_nt_kernel(GemmArgs g) {
  const int total = g.ntiles * g.nbatch * g.nsplit;
  int w = static_cast<int>(blockIdx.x);
  if (g.xcd_remap) w = remap_xcd(w, total);
  gemm_nt_tile<TI, TO, WM, WN, CPLX>(g, w);
}

}  // namespace rr

namespace rr {

// ------------------------------------------------------------------------------------------------
// LDS-staged NT GEMM for the hot real contractions (exchange half-transform and rank-k update).
// A workgroup of 4 waves (2 x 2) computes a BM x BN tile; per k-stage of BK elements the A and B
// panels are loaded from global memory in coalesced 16-byte vectors into registers, written to a
// double-buffered LDS ring (rows padded by 16 bytes), and consumed by the waves as matrix-core
// fragments. The global loads of stage s+1 are in flight while stage s is computed; one barrier
// per stage. Supports batches, split-K partials, lower-triangle tiles and the second output.
// ------------------------------------------------------------------------------------------------
template <typename TI, typename TO, int BM, int BN, int BK, int WGM, int NBUF = 2>
__device__ __forceinline__ void gemm_nt_lds_tile(const GemmArgs& g, int w) {
  static_assert(NBUF == 1 || NBUF == 2, "one or two LDS stage buffers");
  using MM = Mma<TI>;
  using acc_t = typename MM::acc_t;
  using frag = typename MM::frag;
  constexpr int KL = MM::KL;
  constexpr int SUB = 4 * KL;                  // k per fragment sub-step
  static_assert(BK % SUB == 0, "BK must be a multiple of the fragment k-slab");
  constexpr int WGN = 4 / WGM;                 // waves along N
  constexpr int WM = BM / (16 * WGM), WN = BN / (16 * WGN);  // 16x16 tiles per wave
  static_assert(WM * 16 * WGM == BM && WN * 16 * WGN == BN, "tile not divisible over the waves");
  constexpr int VE = 16 / sizeof(TI);          // elements per 16-byte vector
  constexpr int PAD = VE;                      // row padding (16 bytes)
  constexpr int LDR = BK + PAD;                // LDS row stride (elements)
  constexpr int VPR = BK / VE;                 // vectors per row per stage
  constexpr int AV = (BM * VPR + 255) / 256;    // A vectors per thread per stage
  constexpr int BV = (BN * VPR + 255) / 256;
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));

  __shared__ __attribute__((aligned(16))) TI lds[NBUF][(BM + BN) * LDR];

  const int tile = w % g.ntiles;
  const int rest = w / g.ntiles;
  const int batch = rest % g.nbatch;
  const int split = rest / g.nbatch;
  int tm, tn;
  if (g.tri) {
    int t = tile;
    int a = static_cast<int>((sqrtf(8.0f * static_cast<float>(t) + 1.0f) - 1.0f) * 0.5f);
    while ((a + 1) * (a + 2) / 2 <= t) ++a;
    while (a * (a + 1) / 2 > t) --a;
    tm = a;
    tn = t - a * (a + 1) / 2;
  } else {
    tm = tile % g.tiles_m;
    tn = tile / g.tiles_m;
  }

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, grp = lane >> 4;
  const int wm0 = (wave % WGM) * (16 * WM), wn0 = (wave / WGM) * (16 * WN);
  const int bm0 = tm * BM, bn0 = tn * BN;

  const long long kbeg = static_cast<long long>(split) * g.kper;
  long long kend = kbeg + g.kper;
  if (kend > g.K) kend = g.K;

  const TI* Ab = static_cast<const TI*>(g.A) + static_cast<long long>(batch) * g.sAb;
  const TI* Bb = static_cast<const TI*>(g.B) + static_cast<long long>(batch) * g.sBb;

  // per-thread global vector slots: vector v of the stage -> (row, col-vector)
  const TI* asrc[AV];
  int arow_l[AV], acol[AV];
#pragma unroll
  for (int v = 0; v < AV; ++v) {
    int id = tid + v * 256;
    const bool ok = id < BM * VPR;
    if (!ok) id = 0;
    arow_l[v] = ok ? id / VPR : -1;
    acol[v] = (id % VPR) * VE;
    const int m = bm0 + id / VPR;
    asrc[v] = (ok && m < g.M) ? Ab + static_cast<long long>(m) * g.lda : nullptr;
  }
  const TI* bsrc[BV];
  int brow_l[BV], bcol[BV];
#pragma unroll
  for (int v = 0; v < BV; ++v) {
    int id = tid + v * 256;
    const bool ok = id < BN * VPR;
    if (!ok) id = 0;
    brow_l[v] = ok ? id / VPR : -1;
    bcol[v] = (id % VPR) * VE;
    const int n = bn0 + id / VPR;
    bsrc[v] = (ok && n < g.N) ? Bb + static_cast<long long>(n) * g.ldb : nullptr;
  }

  // Stage s+1 is requested from global memory while stage s is computed from LDS.
  u4 ra[AV], rb[BV];
  auto gload = [&](long long k0) {
#pragma unroll
    for (int v = 0; v < AV; ++v) {
      const long long k = k0 + acol[v];
      ra[v] = (asrc[v] && k < kend) ? *reinterpret_cast<const u4*>(asrc[v] + k) : u4{0, 0, 0, 0};
    }
#pragma unroll
    for (int v = 0; v < BV; ++v) {
      const long long k = k0 + bcol[v];
      rb[v] = (bsrc[v] && k < kend) ? *reinterpret_cast<const u4*>(bsrc[v] + k) : u4{0, 0, 0, 0};
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int v = 0; v < AV; ++v)
      if (arow_l[v] >= 0) *reinterpret_cast<u4*>(&lds[buf][arow_l[v] * LDR + acol[v]]) = ra[v];
#pragma unroll
    for (int v = 0; v < BV; ++v)
      if (brow_l[v] >= 0) *reinterpret_cast<u4*>(&lds[buf][(BM + brow_l[v]) * LDR + bcol[v]]) = rb[v];
  };

  acc_t acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = acc_t{};

  const int nstage = static_cast<int>((kend - kbeg + BK - 1) / BK);
  if (nstage > 0) {
    gload(kbeg);
    lstore(0);
  }
  __syncthreads();
  for (int s = 0; s < nstage; ++s) {
    const int buf = NBUF == 2 ? (s & 1) : 0;
    if (s + 1 < nstage) gload(kbeg + static_cast<long long>(s + 1) * BK);
    const TI* la = &lds[buf][0];
    const TI* lb = &lds[buf][BM * LDR];
#pragma unroll
    for (int ks = 0; ks < BK; ks += SUB) {
      frag fa[WM], fb[WN];
#pragma unroll
      for (int i = 0; i < WM; ++i) MM::load(fa[i], la + (wm0 + 16 * i + r) * LDR + ks + grp * KL);
#pragma unroll
      for (int j = 0; j < WN; ++j) MM::load(fb[j], lb + (wn0 + 16 * j + r) * LDR + ks + grp * KL);
      MM::mma_tiles(acc, fa, fb);
    }
    if constexpr (NBUF == 2) {
      if (s + 1 < nstage) lstore(buf ^ 1);
      __syncthreads();
    } else {
      // one buffer: every wave has finished reading stage s before it is overwritten
      __syncthreads();
      if (s + 1 < nstage) {
        lstore(0);
        __syncthreads();
      }
    }
  }

  const long long coff = static_cast<long long>(batch) * g.sCb + static_cast<long long>(split) * g.sCz;
  TO* Cb = static_cast<TO*>(g.C) + coff;
  TO* Cb2 = g.C2 ? static_cast<TO*>(g.C2) + coff : nullptr;
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int n = bn0 + wn0 + 16 * j + r;
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = bm0 + wm0 + 16 * i + MM::out_row(grp, e);
        if (m >= g.M) continue;
        int mm = m, nn = n;
        TO* base = Cb;
        if (g.mcut > 0 && mm >= g.mcut) mm -= g.mcut, base = Cb2;
        if (g.ncut > 0 && nn >= g.ncut) nn -= g.ncut, base = Cb2;
        if (g.nblk > 0) {
          const int b = nn / g.nblk;
          nn -= b * g.nblk;
          base += static_cast<long long>(b) * g.sCnb;
        }
        if (g.mrow > 0) {
          const int b = mm / g.mrow;
          mm -= b * g.mrow;
          base += static_cast<long long>(b) * g.sCmb;
        }
        TO* p = base + static_cast<long long>(mm) * g.ldcm + static_cast<long long>(nn) * g.ldcn;
        double v = g.alpha * static_cast<double>(acc[i][j][e]);
        if (g.beta != 0.0) v += g.beta * Conv<TO>::from(p[0]);
        store_out<TO>(p, v);
      }
    }
}

template <typename TI, typename TO, int BM, int BN, int BK, int WGM, int NBUF = 2>
__global__ __launch_bounds__(256) void gemm_nt_lds_kernel(GemmArgs g) {
  const int total = g.ntiles * g.nbatch * g.nsplit;
  int w = static_cast<int>(blockIdx.x);
  if (g.xcd_remap) w = remap_xcd(w, total);
  gemm_nt_lds_tile<TI, TO, BM, BN, BK, WGM, NBUF>(g, w);
}


// ------------------------------------------------------------------------------------------------
// Multi-stage bf16 NT GEMM fed by direct global->LDS loads (global_load_lds_dwordx4).
//
// Each stage stages BK = 32 k-elements (64 bytes) of BM A-rows and BN B-rows. A DMA instruction moves
// 64 lanes x 16 bytes = 16 rows into LDS (lane l -> row l/4, 16-byte slot l%4); the slot holds the
// global chunk slot ^ ((row/2) % 4), so each 8-lane phase of a fragment load (8 consecutive rows,
// one chunk) covers all 32 LDS banks once. Chunks beyond the rows or beyond the k-range are read from a zero page. NST stages are in
// flight: before computing stage s each wave waits (vmcnt) only for its own loads of stage s, then
// one barrier publishes the stage and frees the buffer of stage s-1 for stage s+NST-1. The DMA is
// issued from inline asm so the compiler adds no conservative wait of its own before the LDS reads;
// the instruction has no register results. Summation order equals the other NT kernels.
// ------------------------------------------------------------------------------------------------
__device__ inline unsigned lds_addr(const void* p) {
  return static_cast<unsigned>(reinterpret_cast<size_t>((const __attribute__((address_space(3))) unsigned char*)p));
}

__device__ inline void lds_dma16(const void* gptr, unsigned lds_uniform) {
  const unsigned m0v = __builtin_amdgcn_readfirstlane(lds_uniform);  // wave-uniform by construction
  unsigned saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %2\n\t"
      "global_load_lds_dwordx4 %1, off\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(saved)
      : "v"(gptr), "s"(m0v)
      : "memory");
}

// s_waitcnt vmcnt(n) for a wave-uniform n (0..15)
__device__ inline void wait_vmcnt(int n) {
  switch (n) {
    case 0: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    case 1: asm volatile("s_waitcnt vmcnt(1)" ::: "memory"); break;
    case 2: asm volatile("s_waitcnt vmcnt(2)" ::: "memory"); break;
    case 3: asm volatile("s_waitcnt vmcnt(3)" ::: "memory"); break;
    case 4: asm volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
    case 5: asm volatile("s_waitcnt vmcnt(5)" ::: "memory"); break;
    case 6: asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); break;
    case 7: asm volatile("s_waitcnt vmcnt(7)" ::: "memory"); break;
    case 8: asm volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
    case 9: asm volatile("s_waitcnt vmcnt(9)" ::: "memory"); break;
    case 10: asm volatile("s_waitcnt vmcnt(10)" ::: "memory"); break;
    case 11: asm volatile("s_waitcnt vmcnt(11)" ::: "memory"); break;
    case 12: asm volatile("s_waitcnt vmcnt(12)" ::: "memory"); break;
    case 13: asm volatile("s_waitcnt vmcnt(13)" ::: "memory"); break;
    case 14: asm volatile("s_waitcnt vmcnt(14)" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(15)" ::: "memory"); break;
  }
}

template <typename TO, int BM, int BN, int WGM, int NST>
__global__ __launch_bounds__(256) void gemm_nt_dma_kernel(GemmArgs g) {
  using MM = Mma<bf16_t>;
  using acc_t = typename MM::acc_t;
  using frag = typename MM::frag;
  constexpr int BK = 32;           // k-elements per stage (one 16x16x32 step)
  constexpr int ROWB = 64;         // bytes per staged row
  constexpr int TI = (BM + BN) / 16;  // DMA instructions per stage
  constexpr int STAGE = (BM + BN) * ROWB;
  constexpr int MAXJ = (TI + 3) / 4;
  constexpr int WGN = 4 / WGM;
  constexpr int WM = BM / (16 * WGM), WN = BN / (16 * WGN);
  static_assert(BM % 16 == 0 && BN % 16 == 0, "tiles are made of 16-row DMA groups");
  static_assert(WM * 16 * WGM == BM && WN * 16 * WGN == BN, "tile not divisible over the waves");
  static_assert(NST >= 2 && (MAXJ * (NST - 2)) <= 15, "wait counts must fit the helper");

  __shared__ __attribute__((aligned(1024))) unsigned char lds[NST * STAGE];

  const int total = g.ntiles * g.nbatch * g.nsplit;
  int w = static_cast<int>(blockIdx.x);
  if (g.xcd_remap) w = remap_xcd(w, total);
  const int tile = w % g.ntiles;
  const int rest = w / g.ntiles;
  const int batch = rest % g.nbatch;
  const int split = rest / g.nbatch;
  int tm, tn;
  if (g.tri) {
    int t = tile;
    int a = static_cast<int>((sqrtf(8.0f * static_cast<float>(t) + 1.0f) - 1.0f) * 0.5f);
    while ((a + 1) * (a + 2) / 2 <= t) ++a;
    while (a * (a + 1) / 2 > t) --a;
    tm = a;
    tn = t - a * (a + 1) / 2;
  } else {
    tm = tile % g.tiles_m;
    tn = tile / g.tiles_m;
  }

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: keeps the stage logic scalar
  const int r = lane & 15, grp = lane >> 4;
  const int wm0 = (wave % WGM) * (16 * WM), wn0 = (wave / WGM) * (16 * WN);
  const int bm0 = tm * BM, bn0 = tn * BN;

  const long long kbeg = static_cast<long long>(split) * g.kper;
  long long kend = kbeg + g.kper;
  if (kend > g.K) kend = g.K;

  const bf16_t* Ab = static_cast<const bf16_t*>(g.A) + static_cast<long long>(batch) * g.sAb;
  const bf16_t* Bb = static_cast<const bf16_t*>(g.B) + static_cast<long long>(batch) * g.sBb;

  // DMA sources of this wave: instruction i = wave + 4 j covers staged rows [16 i, 16 i + 16)
  const int cnt = (TI - wave + 3) / 4;  // wave-uniform
  const bf16_t* src[MAXJ];
  int kch[MAXJ];
#pragma unroll
  for (int j = 0; j < MAXJ; ++j) {
    const int i = wave + 4 * j;
    src[j] = nullptr;
    kch[j] = 0;
    if (i < TI) {
      const int row = 16 * i + (lane >> 2);
      const int trow = row < BM ? row : row - BM;
      const int c = (lane & 3) ^ ((trow >> 1) & 3);
      kch[j] = 8 * c;
      if (row < BM) {
        const int m = bm0 + row;
        if (m < g.M) src[j] = Ab + static_cast<long long>(m) * g.lda + 8 * c;
      } else {
        const int n = bn0 + trow;
        if (n < g.N) src[j] = Bb + static_cast<long long>(n) * g.ldb + 8 * c;
      }
    }
  }
  const unsigned lbase = __builtin_amdgcn_readfirstlane(lds_addr(lds));
  auto issue = [&](int st) {
    const long long k0 = kbeg + static_cast<long long>(st) * BK;
    const unsigned sb = lbase + static_cast<unsigned>((st % NST) * STAGE);
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) {
      if (j < cnt) {
        const void* gp = (src[j] && k0 + kch[j] < kend) ? static_cast<const void*>(src[j] + k0) : g.zpage;
        lds_dma16(gp, sb + static_cast<unsigned>((wave + 4 * j) * 1024));
      }
    }
  };

  acc_t acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = acc_t{};

  const int nstage = static_cast<int>((kend - kbeg + BK - 1) / BK);
  for (int st = 0; st < NST - 1 && st < nstage; ++st) issue(st);
  for (int st = 0; st < nstage; ++st) {
    const int ahead = min(NST - 2, nstage - 1 - st);  // own stages issued after st
    wait_vmcnt(cnt * ahead);
    __builtin_amdgcn_s_barrier();
    if (st + NST - 1 < nstage) issue(st + NST - 1);
    const unsigned char* base = lds + (st % NST) * STAGE;
    frag fa[WM], fb[WN];
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int row = wm0 + 16 * i + r;
      MM::load(fa[i], reinterpret_cast<const bf16_t*>(base + row * ROWB + ((grp ^ ((row >> 1) & 3)) << 4)));
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int row = wn0 + 16 * j + r;
      MM::load(fb[j], reinterpret_cast<const bf16_t*>(base + (BM + row) * ROWB + ((grp ^ ((row >> 1) & 3)) << 4)));
    }
    MM::mma_tiles(acc, fa, fb);
  }

  const long long coff = static_cast<long long>(batch) * g.sCb + static_cast<long long>(split) * g.sCz;
  TO* Cb = static_cast<TO*>(g.C) + coff;
  TO* Cb2 = g.C2 ? static_cast<TO*>(g.C2) + coff : nullptr;
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int n = bn0 + wn0 + 16 * j + r;
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = bm0 + wm0 + 16 * i + MM::out_row(grp, e);
        if (m >= g.M) continue;
        int mm = m, nn = n;
        TO* base = Cb;
        if (g.mcut > 0 && mm >= g.mcut) mm -= g.mcut, base = Cb2;
        if (g.ncut > 0 && nn >= g.ncut) nn -= g.ncut, base = Cb2;
        if (g.nblk > 0) {
          const int b = nn / g.nblk;
          nn -= b * g.nblk;
          base += static_cast<long long>(b) * g.sCnb;
        }
        TO* p = base + static_cast<long long>(mm) * g.ldcm + static_cast<long long>(nn) * g.ldcn;
        double v = g.alpha * static_cast<double>(acc[i][j][e]);
        if (g.beta != 0.0) v += g.beta * Conv<TO>::from(p[0]);
        store_out<TO>(p, v);
      }
    }
}


// Leaner variant of gemm_nt_dma_kernel with the per-stage bookkeeping taken off the hot path:
//  * rows beyond M (N) are clamped to row M-1 (N-1): their products land in accumulators the
//    epilogue never stores, so only the k-range needs the zero page, and only in a partial stage;
//  * a full stage issues each DMA from its precomputed row pointer with one 64-bit add;
//  * the wait before a stage is a compile-time count per wave class (waves holding TI/4 + 1 or TI/4
//    DMA slots), and vmcnt(0) in the last NST-2 stages.
// The staged data, LDS layout and MFMA sequence are those of gemm_nt_dma_kernel (bitwise equal).
template <int N>
__device__ inline void wait_vm() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

template <typename TO, int BM, int BN, int WGM, int NST, int NW = 4>
__global__ __launch_bounds__(64 * NW) void gemm_nt_dma2_kernel(GemmArgs g) {
  using MM = Mma<bf16_t>;
  using acc_t = typename MM::acc_t;
  using frag = typename MM::frag;
  constexpr int BK = 32;
  constexpr int ROWB = 64;
  constexpr int TI = (BM + BN) / 16;
  constexpr int STAGE = (BM + BN) * ROWB;
  constexpr int MAXJ = (TI + NW - 1) / NW;
  constexpr int WGN = NW / WGM;
  constexpr int WM = BM / (16 * WGM), WN = BN / (16 * WGN);
  constexpr int CHI = MAXJ, CLO = TI / NW;  // DMA slots of waves 0..TI%NW-1 and of the others
  static_assert(BM % 16 == 0 && BN % 16 == 0, "tiles are made of 16-row DMA groups");
  static_assert(WGM * WGN == NW && WM * 16 * WGM == BM && WN * 16 * WGN == BN, "tile not divisible over the waves");
  static_assert(NST >= 2 && CHI * (NST - 2) <= 63, "wait counts must fit vmcnt");

  __shared__ __attribute__((aligned(1024))) unsigned char lds[NST * STAGE];

  const int total = g.ntiles * g.nbatch * g.nsplit;
  int w = static_cast<int>(blockIdx.x);
  if (g.xcd_remap) w = remap_xcd(w, total);
  const int tile = w % g.ntiles;
  const int rest = w / g.ntiles;
  const int batch = rest % g.nbatch;
  const int split = rest / g.nbatch;
  int tm, tn;
  if (g.tri) {
    int t = tile;
    int a = static_cast<int>((sqrtf(8.0f * static_cast<float>(t) + 1.0f) - 1.0f) * 0.5f);
    while ((a + 1) * (a + 2) / 2 <= t) ++a;
    while (a * (a + 1) / 2 > t) --a;
    tm = a;
    tn = t - a * (a + 1) / 2;
  } else {
    tm = tile % g.tiles_m;
    tn = tile / g.tiles_m;
  }

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, grp = lane >> 4;
  const int wm0 = (wave % WGM) * (16 * WM), wn0 = (wave / WGM) * (16 * WN);
  const int bm0 = tm * BM, bn0 = tn * BN;

  const long long kbeg = static_cast<long long>(split) * g.kper;
  long long kend = kbeg + g.kper;
  if (kend > g.K) kend = g.K;

  const bf16_t* Ab = static_cast<const bf16_t*>(g.A) + static_cast<long long>(batch) * g.sAb;
  const bf16_t* Bb = static_cast<const bf16_t*>(g.B) + static_cast<long long>(batch) * g.sBb;

  const bool hi = wave < (TI % NW) || (TI % NW) == 0;  // this wave holds CHI slots (else CLO)
  const bf16_t* src[MAXJ];
  int kch[MAXJ];
#pragma unroll
  for (int j = 0; j < MAXJ; ++j) {
    const int i = wave + NW * j;
    const int row = 16 * min(i, TI - 1) + (lane >> 2);
    const int trow = row < BM ? row : row - BM;
    const int c = (lane & 3) ^ ((trow >> 1) & 3);
    kch[j] = 8 * c;
    if (row < BM) {
      const int m = min(bm0 + row, g.M - 1);
      src[j] = Ab + static_cast<long long>(m) * g.lda + 8 * c + kbeg;
    } else {
      const int n = min(bn0 + trow, g.N - 1);
      src[j] = Bb + static_cast<long long>(n) * g.ldb + 8 * c + kbeg;
    }
  }
  const unsigned lbase = __builtin_amdgcn_readfirstlane(lds_addr(lds));
  const int nstage = static_cast<int>((kend - kbeg + BK - 1) / BK);
  const int nfull = static_cast<int>((kend - kbeg) / BK);
  auto issue = [&](int st) {
    const unsigned sb = lbase + static_cast<unsigned>((st % NST) * STAGE);
    const long long koff = static_cast<long long>(st) * BK;
    if (st < nfull) {
#pragma unroll
      for (int j = 0; j < MAXJ; ++j)
        if (j < CLO || (j < CHI && hi)) lds_dma16(src[j] + koff, sb + static_cast<unsigned>((wave + NW * j) * 1024));
    } else {
#pragma unroll
      for (int j = 0; j < MAXJ; ++j)
        if (j < CLO || (j < CHI && hi)) {
          const void* gp = (kbeg + koff + kch[j] < kend) ? static_cast<const void*>(src[j] + koff) : g.zpage;
          lds_dma16(gp, sb + static_cast<unsigned>((wave + NW * j) * 1024));
        }
    }
  };

  acc_t acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = acc_t{};

  for (int st = 0; st < NST - 1 && st < nstage; ++st) issue(st);
  // Retire the kernel-argument scalar loads here: with none pending inside the loop, the compiler can
  // wait for the fragment reads one by one (in-order LDS returns) instead of for all of them.
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt / expcnt untouched
  for (int st = 0; st < nstage; ++st) {
    if (st + NST - 2 < nstage) {  // NST-2 later stages of this wave are in flight
      if (hi)
        wait_vm<CHI * (NST - 2)>();
      else
        wait_vm<CLO * (NST - 2)>();
    } else {
      wait_vm<0>();
    }
    __builtin_amdgcn_s_barrier();
    if (st + NST - 1 < nstage) issue(st + NST - 1);
    const unsigned char* base = lds + (st % NST) * STAGE;
    frag fa[WM], fb[WN];
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int row = wm0 + 16 * i + r;
      MM::load(fa[i], reinterpret_cast<const bf16_t*>(base + row * ROWB + ((grp ^ ((row >> 1) & 3)) << 4)));
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int row = wn0 + 16 * j + r;
      MM::load(fb[j], reinterpret_cast<const bf16_t*>(base + (BM + row) * ROWB + ((grp ^ ((row >> 1) & 3)) << 4)));
    }
    MM::mma_tiles(acc, fa, fb);
  }

  const long long coff = static_cast<long long>(batch) * g.sCb + static_cast<long long>(split) * g.sCz;
  TO* Cb = static_cast<TO*>(g.C) + coff;
  TO* Cb2 = g.C2 ? static_cast<TO*>(g.C2) + coff : nullptr;
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int n = bn0 + wn0 + 16 * j + r;
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = bm0 + wm0 + 16 * i + MM::out_row(grp, e);
        if (m >= g.M) continue;
        int mm = m, nn = n;
        TO* base = Cb;
        if (g.mcut > 0 && mm >= g.mcut) mm -= g.mcut, base = Cb2;
        if (g.ncut > 0 && nn >= g.ncut) nn -= g.ncut, base = Cb2;
        if (g.nblk > 0) {
          const int b = nn / g.nblk;
          nn -= b * g.nblk;
          base += static_cast<long long>(b) * g.sCnb;
        }
        TO* p = base + static_cast<long long>(mm) * g.ldcm + static_cast<long long>(nn) * g.ldcn;
        double v = g.alpha * static_cast<double>(acc[i][j][e]);
        if (g.beta != 0.0) v += g.beta * Conv<TO>::from(p[0]);
        store_out<TO>(p, v);
      }
    }
}


// ------------------------------------------------------------------------------------------------
// FP64 / FP32 direct-to-LDS GEMM with NW waves per workgroup (two per SIMD at NW = 8, so one wave's
// stalls are covered by the other's matrix-core work even with one workgroup per CU). A stage holds
// one 128-byte line (16 doubles or 32 floats) of each of the BM + BN rows; a DMA instruction moves 8
// rows (lane l: row l / 8, 16-byte slot l % 8 holding chunk slot ^ (row % 8), so the eight rows of a
// fragment phase hit all 32 banks). Fragments and matrix-core steps are those of Mma<TI> (lane group
// g takes the 32 bytes 32 g .. 32 g + 31 of the stage), so sums are bitwise equal to the other
// kernels of the operand type.
// Row clamping, compile-time waits and the zero page for a partial stage as in gemm_nt_dma2_kernel.
// ------------------------------------------------------------------------------------------------
// 16-byte slot swizzle of a 128-byte staged row: the slot of chunk c is c ^ line_swizzle(row). A
// 16-byte fragment read covers 16 consecutive rows in one pass of the 64 LDS banks: with (row mod 8)
// alone, rows r and r + 8 land on the same banks (half of the LDS cycles were bank conflicts,
// SQ_LDS_BANK_CONFLICT); flipping the low bit for the second group of eight rows separates them.
__device__ inline int line_swizzle(int row) { return (row & 7) ^ ((row >> 3) & 1); }

template <typename TI, typename TO, int BM, int BN, int WGM, int NST, int NW>
__global__ __launch_bounds__(64 * NW) void gemm_nt_dma64_kernel(GemmArgs g) {
  constexpr bool B16 = std::is_same<TI, bf16_t>::value;  // two 32-k matrix-core steps per stage
  using MM = Mma<TI>;
  using acc_t = typename MM::acc_t;
  using frag = typename MM::frag;
  constexpr int BK = 128 / sizeof(TI);  // elements per staged row
  constexpr int VE = 16 / sizeof(TI);    // elements per 16-byte chunk
  constexpr int ROWB = 128;
  constexpr int NDMA = (BM + BN) / 8;  // DMA instructions per stage
  constexpr int STAGE = (BM + BN) * ROWB;
  constexpr int MAXJ = (NDMA + NW - 1) / NW;
  constexpr int WGN = NW / WGM;
  constexpr int WM = BM / (16 * WGM), WN = BN / (16 * WGN);
  constexpr int CHI = MAXJ, CLO = NDMA / NW;
  static_assert(BM % 16 == 0 && BN % 16 == 0, "tiles are made of 16-row fragments");
  static_assert(WGM * WGN == NW && WM * 16 * WGM == BM && WN * 16 * WGN == BN, "tile not divisible over the waves");
  static_assert(NST >= 2 && CHI * (NST - 2) <= 63, "wait counts must fit vmcnt");

  __shared__ __attribute__((aligned(1024))) unsigned char lds[NST * STAGE];

  const int total = g.ntiles * g.nbatch * g.nsplit;
  int w = static_cast<int>(blockIdx.x);
  if (g.xcd_remap) w = remap_xcd(w, total);
  const int tile = w % g.ntiles;
  const int rest = w / g.ntiles;
  const int batch = rest % g.nbatch;
  const int split = rest / g.nbatch;
  int tm, tn;
  if (g.tri) {
    int t = tile;
    int a = static_cast<int>((sqrtf(8.0f * static_cast<float>(t) + 1.0f) - 1.0f) * 0.5f);
    while ((a + 1) * (a + 2) / 2 <= t) ++a;
    while (a * (a + 1) / 2 > t) --a;
    tm = a;
    tn = t - a * (a + 1) / 2;
  } else {
    tm = tile % g.tiles_m;
    tn = tile / g.tiles_m;
  }

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, grp = lane >> 4;
  const int wm0 = (wave % WGM) * (16 * WM), wn0 = (wave / WGM) * (16 * WN);
  const int bm0 = tm * BM, bn0 = tn * BN;

  const long long kbeg = static_cast<long long>(split) * g.kper;
  long long kend = kbeg + g.kper;
  if (kend > g.K) kend = g.K;

  const TI* Ab = static_cast<const TI*>(g.A) + static_cast<long long>(batch) * g.sAb;
  const TI* Bb = static_cast<const TI*>(g.B) + static_cast<long long>(batch) * g.sBb;

  const bool hi = wave < (NDMA % NW) || (NDMA % NW) == 0;
  const TI* src[MAXJ];
  int kch[MAXJ];
#pragma unroll
  for (int j = 0; j < MAXJ; ++j) {
    const int i = wave + NW * j;
    const int row = 8 * min(i, NDMA - 1) + (lane >> 3);
    const int c = (lane & 7) ^ line_swizzle(row);
    kch[j] = VE * c;
    if (row < BM) {
      const int m = min(bm0 + row, g.M - 1);
      src[j] = Ab + static_cast<long long>(m) * g.lda + VE * c + kbeg;
    } else {
      const int n = min(bn0 + row - BM, g.N - 1);
      src[j] = Bb + static_cast<long long>(n) * g.ldb + VE * c + kbeg;
    }
  }
  const unsigned lbase = __builtin_amdgcn_readfirstlane(lds_addr(lds));
  const int nstage = static_cast<int>((kend - kbeg + BK - 1) / BK);
  const int nfull = static_cast<int>((kend - kbeg) / BK);
  auto issue = [&](int st) {
    const unsigned sb = lbase + static_cast<unsigned>((st % NST) * STAGE);
    const long long koff = static_cast<long long>(st) * BK;
    if (st < nfull) {
#pragma unroll
      for (int j = 0; j < MAXJ; ++j)
        if (j < CLO || (j < CHI && hi)) lds_dma16(src[j] + koff, sb + static_cast<unsigned>((wave + NW * j) * 1024));
    } else {
#pragma unroll
      for (int j = 0; j < MAXJ; ++j)
        if (j < CLO || (j < CHI && hi)) {
          const void* gp = (kbeg + koff + kch[j] < kend) ? static_cast<const void*>(src[j] + koff) : g.zpage;
          lds_dma16(gp, sb + static_cast<unsigned>((wave + NW * j) * 1024));
        }
    }
  };

  acc_t acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = acc_t{};

  for (int st = 0; st < NST - 1 && st < nstage; ++st) issue(st);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): no scalar loads pending inside the loop
  for (int st = 0; st < nstage; ++st) {
    if (st + NST - 2 < nstage) {
      if (hi)
        wait_vm<CHI * (NST - 2)>();
      else
        wait_vm<CLO * (NST - 2)>();
    } else {
      wait_vm<0>();
    }
    __builtin_amdgcn_s_barrier();
    if (st + NST - 1 < nstage) issue(st + NST - 1);
    const unsigned char* base = lds + (st % NST) * STAGE;
    frag fa[WM], fb[WN];
    if constexpr (B16) {
      // step h reads chunk 4 h + grp: lane group g holds k = 32 h + 8 g .. + 7, the k order of the
      // other bf16 kernels
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        frag ga[WM], gb[WN];
#pragma unroll
        for (int i = 0; i < WM; ++i) {
          const int lrow = wm0 + 16 * i + r;
          MM::load(ga[i], reinterpret_cast<const TI*>(base + lrow * ROWB + (((4 * h + grp) ^ line_swizzle(lrow)) << 4)));
        }
#pragma unroll
        for (int j = 0; j < WN; ++j) {
          const int lrow = BM + wn0 + 16 * j + r;
          MM::load(gb[j], reinterpret_cast<const TI*>(base + lrow * ROWB + (((4 * h + grp) ^ line_swizzle(lrow)) << 4)));
        }
        MM::mma_tiles(acc, ga, gb);
      }
      continue;
    }
    auto lfrag = [&](frag& f, int lrow) {
      const unsigned char* rp = base + lrow * ROWB;
      const int x = line_swizzle(lrow);
      if constexpr (sizeof(TI) == 8) {
        const d2 a = *reinterpret_cast<const d2*>(rp + (((2 * grp) ^ x) << 4));
        const d2 b = *reinterpret_cast<const d2*>(rp + (((2 * grp + 1) ^ x) << 4));
        f.v[0] = a.x, f.v[1] = a.y, f.v[2] = b.x, f.v[3] = b.y;
      } else if constexpr (sizeof(TI) == 4) {
        const f4 a = *reinterpret_cast<const f4*>(rp + (((2 * grp) ^ x) << 4));
        const f4 b = *reinterpret_cast<const f4*>(rp + (((2 * grp + 1) ^ x) << 4));
        f.v[0] = a.x, f.v[1] = a.y, f.v[2] = a.z, f.v[3] = a.w;
        f.v[4] = b.x, f.v[5] = b.y, f.v[6] = b.z, f.v[7] = b.w;
      }
    };
#pragma unroll
    for (int i = 0; i < WM; ++i) lfrag(fa[i], wm0 + 16 * i + r);
#pragma unroll
    for (int j = 0; j < WN; ++j) lfrag(fb[j], BM + wn0 + 16 * j + r);
    MM::mma_tiles(acc, fa, fb);
  }

  const long long coff = static_cast<long long>(batch) * g.sCb + static_cast<long long>(split) * g.sCz;
  TO* Cb = static_cast<TO*>(g.C) + coff;
  TO* Cb2 = g.C2 ? static_cast<TO*>(g.C2) + coff : nullptr;
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int n = bn0 + wn0 + 16 * j + r;
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = bm0 + wm0 + 16 * i + MM::out_row(grp, e);
        if (m >= g.M) continue;
        int mm = m, nn = n;
        TO* basep = Cb;
        if (g.mcut > 0 && mm >= g.mcut) mm -= g.mcut, basep = Cb2;
        if (g.ncut > 0 && nn >= g.ncut) nn -= g.ncut, basep = Cb2;
        if (g.nblk > 0) {
          const int b = nn / g.nblk;
          nn -= b * g.nblk;
          basep += static_cast<long long>(b) * g.sCnb;
        }
        TO* p = basep + static_cast<long long>(mm) * g.ldcm + static_cast<long long>(nn) * g.ldcn;
        double v = g.alpha * static_cast<double>(acc[i][j][e]);
        if (g.beta != 0.0) v += g.beta * Conv<TO>::from(p[0]);
        store_out<TO>(p, v);
      }
    }
}

}  // namespace rr
