// Row-merged bf16 half-transform GEMM: C[(b, m)][n] = sum_k A_b[m][k] B[n][k] for every batch b of
// A (the 3-index tensor) against one shared operand B (the packed factor rows), computed as a single
// tall product over the merged rows (b, m) so that no tile is cut at a batch boundary.
//
// Design for one CDNA4 compute unit (160 KB LDS, 4 SIMDs, 512 VGPRs per SIMD lane):
//  * one persistent workgroup of eight waves per CU (two per SIMD: a single wave per SIMD leaves the
//    matrix cores a third idle), looping over 128-row x BN-column tiles, BN = 16 NT up to 448;
//  * each wave owns 64 rows x (7 or 6) 16-column blocks: the NT blocks are split 4 ways with the
//    wider block of each SIMD's pair on a different column group, so every SIMD gets the same work;
//  * the A rows and the B rows of a 32-k stage are staged by direct global->LDS DMA
//    (global_load_lds_dwordx4) into NST stage buffers; the DMA of stage s+NST-1, the fragment reads
//    of stage s+1 and the matrix-core work of stage s are interleaved column block by column block,
//    so the texture path, the LDS and the matrix cores run side by side;
//  * the matrix-core operands are swapped (B fragment first), so every lane holds four consecutive
//    output columns of one row and stores them as one 8-byte write;
//  * the next tile's first NST-1 stages are issued before the current tile's stores.
// The k order of every output element is that of the other NT kernels (32-element slabs in
// increasing k, one v_mfma_f32_16x16x32_bf16 each), so results are bitwise equal to them.
#pragma once

#include "rr_gemm.hpp"

namespace rr {

// Eligibility of a launch for gemm_rows_bf16_kernel<TO, NT, NST> (checked on the host).
inline bool rows_kernel_fits(const GemmArgs& g, int nt, int nst, size_t out_size) {
  const long long mtot = static_cast<long long>(g.M) * g.nbatch;
  if (g.M <= 0 || g.N <= 0 || g.nbatch <= 0 || mtot >= (1ll << 31)) return false;
  if (g.N > 16 * nt || g.K < 32 * nst || g.K % 8 != 0) return false;
  if (g.tri || g.nsplit != 1 || g.kbatch != 0 || g.mcut != 0 || g.mrow != 0 || g.ldcn != 1) return false;
  if (g.sBb != 0 || g.lda % 8 != 0 || g.sAb % 8 != 0 || g.ldb % 8 != 0) return false;
  if (g.beta != 0.0 || g.alpha != 1.0) return false;
  // 8-byte (4-element) aligned output groups that never straddle a column cut or block
  if (g.N % 4 != 0 || g.ldcm % 4 != 0 || g.sCb % 4 != 0) return false;
  if (g.ncut > 0 && (g.ncut % 4 != 0 || g.C2 == nullptr)) return false;
  if (g.nblk > 0 && (g.nblk % 4 != 0 || g.sCnb % 4 != 0)) return false;
  if (out_size == 2) {  // one 8-byte store per lane and row
    if (reinterpret_cast<uintptr_t>(g.C) % 8 != 0) return false;
    if (g.C2 && reinterpret_cast<uintptr_t>(g.C2) % 8 != 0) return false;
  }
  if (reinterpret_cast<uintptr_t>(g.A) % 16 != 0 || reinterpret_cast<uintptr_t>(g.B) % 16 != 0) return false;
  return true;
}

template <typename TO, int NT, int NST>
__global__ __launch_bounds__(512) void gemm_rows_bf16_kernel(GemmArgs g) {
  constexpr int BM = 128, BN = 16 * NT, BK = 32, ROWB = 64, NW = 8;
  constexpr int TI = (BM + BN) / 16;  // 16-row DMA groups per stage
  constexpr int STAGE = (BM + BN) * ROWB;
  constexpr int CHI = (TI + NW - 1) / NW, CLO = TI / NW;  // DMA slots of waves 0..TI%NW-1 and of the rest
  constexpr int WNMAX = (NT + 3) / 4, WNMIN = NT / 4;
  static_assert(NT >= 13 && NT <= 28 && NST >= 3 && CHI * (NST - 2) <= 63, "tile shape");
  static_assert(NST * STAGE <= 160 * 1024, "stage buffers exceed the LDS");
  __shared__ __attribute__((aligned(1024))) unsigned char lds[NST * STAGE];

  const int mtot = g.M * g.nbatch;  // < 2^31 (host check)
  const int ntiles = g.tiles_m * g.tiles_n;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, grp = lane >> 4;
  const int mi = wave >> 2, ni = ((wave & 3) + mi) & 3;
  const int cnt = WNMIN + (ni < NT % 4 ? 1 : 0);
  const int c0 = ni * WNMIN + (ni < NT % 4 ? ni : NT % 4);
  const int wm0 = 64 * mi, wn0 = 16 * c0;
  const int K = g.K;
  const int nstage = (K + BK - 1) / BK;  // >= NST (host check)
  const bf16_t* A = static_cast<const bf16_t*>(g.A);
  const bf16_t* B = static_cast<const bf16_t*>(g.B);
  const bool hi = wave < (TI % NW) || (TI % NW) == 0;
  const unsigned lbase = __builtin_amdgcn_readfirstlane(lds_addr(lds));

  // (batch, row) of merged row m, for m in [m0, m0 + span) with m0 = b0 * M + r0
  auto split_row = [&](int b0, int r0, int off, int& b, int& rr) {
    b = b0, rr = r0 + off;
    while (rr >= g.M) rr -= g.M, ++b;
  };
  const bf16_t* src[CHI];
  int kch[CHI];
  auto setup = [&](int tile) {
    const int tm = tile % g.tiles_m, tn = tile / g.tiles_m;
    const int bm0 = tm * BM, bn0 = tn * BN;
    const int b0 = bm0 / g.M, r0 = bm0 - b0 * g.M;
#pragma unroll
    for (int j = 0; j < CHI; ++j) {
      const int i = min(wave + NW * j, TI - 1);
      const int row = 16 * i + (lane >> 2);
      const int trow = row < BM ? row : row - BM;
      const int c = (lane & 3) ^ ((trow >> 1) & 3);  // 16-byte chunk swizzle: fragment reads hit all banks
      kch[j] = 8 * c;
      if (row < BM) {
        int b, rr;
        split_row(b0, r0, min(row, mtot - 1 - bm0), b, rr);  // rows past the end repeat the last row
        src[j] = A + b * g.sAb + static_cast<long long>(rr) * g.lda + 8 * c;
      } else {
        const int n = min(bn0 + trow, g.N - 1);
        src[j] = B + static_cast<long long>(n) * g.ldb + 8 * c;
      }
    }
  };
  auto dma = [&](int st, int j) {
    const unsigned sb = lbase + static_cast<unsigned>((st % NST) * STAGE);
    const int koff = st * BK;
    const void* gp = (koff + kch[j] < K) ? static_cast<const void*>(src[j] + koff) : g.zpage;
    lds_dma16(gp, sb + static_cast<unsigned>((wave + NW * j) * 1024));
  };
  auto issue = [&](int st) {
#pragma unroll
    for (int j = 0; j < CHI; ++j)
      if (j < CLO || hi) dma(st, j);
  };
  struct Frags {
    bf8 a[4], b[WNMAX];
  };
  auto read_a = [&](const unsigned char* base, int i) {
    const int row = wm0 + 16 * i + r;
    return *reinterpret_cast<const bf8*>(base + row * ROWB + ((grp ^ ((row >> 1) & 3)) << 4));
  };
  auto read_b = [&](const unsigned char* base, int j) {  // narrow blocks re-read their last fragment
    const int row = wn0 + 16 * (j < WNMIN ? j : min(j, cnt - 1)) + r;
    return *reinterpret_cast<const bf8*>(base + (BM + row) * ROWB + ((grp ^ ((row >> 1) & 3)) << 4));
  };
  f4 acc[4][WNMAX];
  // Matrix-core work of one stage on `cur`, interleaved with the DMA of stage `ist` and the reads of
  // stage `rst` into `nxt` (either skipped when negative).
  auto body = [&](const Frags& cur, Frags& nxt, int ist, int rst) {
    const unsigned char* nbase = lds + ((rst < 0 ? 0 : rst) % NST) * STAGE;
#pragma unroll
    for (int j = 0; j < WNMAX; ++j) {
      if (ist >= 0 && j < CHI && (j < CLO || hi)) dma(ist, j);
      if (rst >= 0) {
        if (j < 4) nxt.a[j] = read_a(nbase, j);
        nxt.b[j] = read_b(nbase, j);
      }
      if (j < WNMIN || cnt > WNMIN) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.b[j], cur.a[i], acc[i][j], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // before the work of stage s: this wave's DMA of stage s+1 complete, stages s+2 .. min(s+NST-2,
  // nstage-1) may stay in flight
  auto tail_wait = [&](int s) {
    const int later = max(0, min(s + NST - 2, nstage - 1) - s - 1);
    wait_vmcnt((hi ? CHI : CLO) * later);
  };

  int tile = blockIdx.x;
  if (tile >= ntiles) return;
  setup(tile);
  for (int st = 0; st < NST - 1; ++st) issue(st);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): kernel arguments in
  TO* C = static_cast<TO*>(g.C);
  TO* C2 = static_cast<TO*>(g.C2);
  while (true) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < WNMAX; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    Frags f0, f1;
    // stage 0 (the stores of the previous tile, issued after these DMAs, are waited for as well)
    if (hi) wait_vm<CHI * (NST - 2)>(); else wait_vm<CLO * (NST - 2)>();
    __builtin_amdgcn_s_barrier();
#pragma unroll
    for (int i = 0; i < 4; ++i) f0.a[i] = read_a(lds, i);
#pragma unroll
    for (int j = 0; j < WNMAX; ++j) f0.b[j] = read_b(lds, j);
    int st = 0;
    for (; st + NST < nstage; st += 2) {  // steady pairs: the DMA of stage st+NST exists
      if (hi) wait_vm<CHI * (NST - 3)>(); else wait_vm<CLO * (NST - 3)>();
      __builtin_amdgcn_s_barrier();  // stage st+1 visible; every wave's reads of stage st-1 done
      body(f0, f1, st + NST - 1, st + 1);
      if (hi) wait_vm<CHI * (NST - 3)>(); else wait_vm<CLO * (NST - 3)>();
      __builtin_amdgcn_s_barrier();
      body(f1, f0, st + NST, st + 2);
    }
    for (; st < nstage; st += 2) {
      tail_wait(st);
      __builtin_amdgcn_s_barrier();
      body(f0, f1, st + NST - 1 < nstage ? st + NST - 1 : -1, st + 1 < nstage ? st + 1 : -1);
      if (st + 1 < nstage) {
        tail_wait(st + 1);
        __builtin_amdgcn_s_barrier();
        body(f1, f0, st + NST < nstage ? st + NST : -1, st + 2 < nstage ? st + 2 : -1);
      }
    }
    // every wave is done with the stage buffers: start the next tile's first stages, then store
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    const int tm = tile % g.tiles_m, tn = tile / g.tiles_m;
    const int next = tile + static_cast<int>(gridDim.x);
    if (next < ntiles) {
      setup(next);
      for (int s2 = 0; s2 < NST - 1; ++s2) issue(s2);
    }
    const int bm0 = tm * BM, bn0 = tn * BN;
    const int b0 = bm0 / g.M, r0 = bm0 - b0 * g.M;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int off = wm0 + 16 * i + r;
      if (bm0 + off >= mtot) continue;
      int b, rr;
      split_row(b0, r0, off, b, rr);
      const long long rowoff = b * g.sCb + static_cast<long long>(rr) * g.ldcm;
#pragma unroll
      for (int j = 0; j < WNMAX; ++j) {
        if (!(j < WNMIN || cnt > WNMIN)) continue;
        int nn = bn0 + wn0 + 16 * j + 4 * grp;
        if (nn >= g.N) continue;
        TO* base = C;
        if (g.ncut > 0 && nn >= g.ncut) nn -= g.ncut, base = C2;
        if (g.nblk > 0) {
          const int blk = nn / g.nblk;
          nn -= blk * g.nblk;
          base += blk * g.sCnb;
        }
        TO* p = base + rowoff + nn;
        if constexpr (std::is_same<TO, bf16_t>::value) {
          const unsigned lo = (unsigned)f32_to_bf16(acc[i][j][0]).bits | ((unsigned)f32_to_bf16(acc[i][j][1]).bits << 16);
          const unsigned hi2 = (unsigned)f32_to_bf16(acc[i][j][2]).bits | ((unsigned)f32_to_bf16(acc[i][j][3]).bits << 16);
          *reinterpret_cast<uint2*>(p) = uint2{lo, hi2};
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) store_out<TO>(p + e, static_cast<double>(acc[i][j][e]));
        }
      }
    }
    if (next >= ntiles) break;
    tile = next;
  }
}

}  // namespace rr

namespace rr {

// Eligibility of a split-K product for gemm_rankk_bf16_kernel (checked on the host).
inline bool rankk_kernel_fits(const GemmArgs& g) {
  if (g.M <= 0 || g.N <= 0 || g.nbatch != 1 || g.tri || g.kbatch != 0) return false;
  if (g.mcut != 0 || g.ncut != 0 || g.mrow != 0 || g.ldcn != 1) return false;
  // column blocks: the four consecutive columns a lane stores lie in one block
  if (g.nblk < 0 || g.nblk % 4 != 0) return false;
  if (g.beta != 0.0 || g.alpha != 1.0 || g.lda % 8 != 0 || g.ldb % 8 != 0 || g.K % 8 != 0) return false;
  if (g.nsplit > 1 && g.kper % 32 != 0) return false;
  const long long cend = g.nblk > 0 ? (g.N - 1) / g.nblk * g.sCnb + static_cast<long long>(g.M) * g.ldcm + g.nblk
                                    : static_cast<long long>(g.M) * g.ldcm + g.N;
  if (cend >= (1ll << 40)) return false;
  if (reinterpret_cast<uintptr_t>(g.A) % 16 != 0 || reinterpret_cast<uintptr_t>(g.B) % 16 != 0) return false;
  return true;
}

// Split-K rank-k update C_z[m][n] = sum_{k in split z} A[m][k] B[n][k] (bf16 operands, fp32 matrix-core
// sums, TO partial outputs): the exchange step K = W_L W_R^T over the flat (aux row, column) index;
// with column blocks (COLBLK, nblk > 0) one launch writes the K of several stacked W_R that share W_L.
// Same building blocks as gemm_rows_bf16_kernel: one persistent workgroup of eight waves per CU,
// 256 x 256 tiles (waves of 128 x 64: eight A and four B fragments for 32 matrix-core steps per
// stage), direct-to-LDS stages, DMA / fragment reads / matrix-core work interleaved, swapped
// operands for four consecutive output columns per lane. Work items (tile, split) are dealt so that
// the workgroups of one XCD take whole splits: the A and B panels of a split are then read through
// that XCD's L2 once for all its tiles. The k order per output equals the other NT kernels' (32-k
// slabs in increasing k inside the split), so partial outputs are bitwise equal to theirs.
template <typename TO, int NST, bool COLBLK>
__global__ __launch_bounds__(512) void gemm_rankk_bf16_kernel(GemmArgs g) {
  constexpr int BM = 256, BN = 256, BK = 32, ROWB = 64, NW = 8;
  constexpr int TI = (BM + BN) / 16;
  constexpr int STAGE = (BM + BN) * ROWB;
  constexpr int CH = TI / NW;  // DMA slots per wave and stage
  constexpr int WM = 8, WN = 4;
  static_assert(TI % NW == 0 && NST >= 3 && CH * (NST - 2) <= 15, "stage shape");
  static_assert(NST * STAGE <= 160 * 1024, "stage buffers exceed the LDS");
  __shared__ __attribute__((aligned(1024))) unsigned char lds[NST * STAGE];

  const int ntiles = g.tiles_m * g.tiles_n;
  const int items = ntiles * g.nsplit;
  const int G = static_cast<int>(gridDim.x);
  const int w = static_cast<int>(blockIdx.x);
  const int lslot = (G % 8 == 0) ? (w % 8) * (G / 8) + w / 8 : w;  // XCD-major slot of this workgroup
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, grp = lane >> 4;
  const int wm0 = 128 * (wave >> 2), wn0 = 64 * (wave & 3);
  const bf16_t* A = static_cast<const bf16_t*>(g.A);
  const bf16_t* B = static_cast<const bf16_t*>(g.B);
  const unsigned lbase = __builtin_amdgcn_readfirstlane(lds_addr(lds));

  int tm = 0, tn = 0, split = 0, kbeg = 0, kend = 0, nstage = 0;
  auto decode = [&](int item) {
    split = item / ntiles;
    const int tile = item - split * ntiles;
    tm = tile % g.tiles_m, tn = tile / g.tiles_m;
    kbeg = split * g.kper;
    kend = min(g.K, kbeg + g.kper);
    nstage = (kend - kbeg + BK - 1) / BK;
  };
  const bf16_t* src[CH];
  int kch[CH];
  auto setup = [&]() {
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int row = 16 * (wave + NW * j) + (lane >> 2);
      const int trow = row < BM ? row : row - BM;
      const int c = (lane & 3) ^ ((trow >> 1) & 3);
      kch[j] = 8 * c;
      if (row < BM) {
        const int m = min(tm * BM + row, g.M - 1);
        src[j] = A + static_cast<long long>(m) * g.lda + 8 * c + kbeg;
      } else {
        const int n = min(tn * BN + trow, g.N - 1);
        src[j] = B + static_cast<long long>(n) * g.ldb + 8 * c + kbeg;
      }
    }
  };
  auto dma = [&](int st, int j) {
    const unsigned sb = lbase + static_cast<unsigned>((st % NST) * STAGE);
    const int koff = st * BK;
    const void* gp = (kbeg + koff + kch[j] < kend) ? static_cast<const void*>(src[j] + koff) : g.zpage;
    lds_dma16(gp, sb + static_cast<unsigned>((wave + NW * j) * 1024));
  };
  auto issue_first = [&]() {
    for (int st = 0; st < NST - 1 && st < nstage; ++st) {
#pragma unroll
      for (int j = 0; j < CH; ++j) dma(st, j);
    }
  };
  auto read_a = [&](const unsigned char* base, int i) {
    const int row = wm0 + 16 * i + r;
    return *reinterpret_cast<const bf8*>(base + row * ROWB + ((grp ^ ((row >> 1) & 3)) << 4));
  };
  auto read_b = [&](const unsigned char* base, int j) {
    const int row = wn0 + 16 * j + r;
    return *reinterpret_cast<const bf8*>(base + (BM + row) * ROWB + ((grp ^ ((row >> 1) & 3)) << 4));
  };
  f4 acc[WM][WN];
  // A fragments are replaced in place by the next stage's right after their last use; B fragments
  // (used by every row group) are double-buffered. Keeps the wave within 256 VGPRs without spills.
  bf8 fa[WM], fb0[WN], fb1[WN];
  auto body = [&](const bf8 (&bcur)[WN], bf8 (&bnxt)[WN], int ist, int rst) {
    const unsigned char* nbase = lds + ((rst < 0 ? 0 : rst) % NST) * STAGE;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
#pragma unroll
      for (int j = 0; j < WN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bcur[j], fa[i], acc[i][j], 0, 0, 0);
      if (rst >= 0) {
        fa[i] = read_a(nbase, i);
        if (i < WN) bnxt[i] = read_b(nbase, i);
      }
      if (ist >= 0 && i < CH) dma(ist, i);
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // this wave's stages issued after the one needed before the work of stage s (s+1 .. min(s+NST-2, nstage-1))
  auto wait_for = [&](int needed) {
    const int later = max(0, min(needed + NST - 3, nstage - 1) - needed);
    wait_vmcnt(CH * later);
  };

  int round = 0;
  int item = lslot;
  if (item >= items) return;
  decode(item);
  setup();
  issue_first();
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): kernel arguments in
  while (true) {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    wait_vmcnt(CH * min(NST - 2, nstage - 1));  // stage 0 in; also retires the previous item's stores
    __builtin_amdgcn_s_barrier();
#pragma unroll
    for (int i = 0; i < WM; ++i) fa[i] = read_a(lds, i);
#pragma unroll
    for (int j = 0; j < WN; ++j) fb0[j] = read_b(lds, j);
    int st = 0;
    for (; st + NST < nstage; st += 2) {
      wait_vm<CH * (NST - 3)>();
      __builtin_amdgcn_s_barrier();
      body(fb0, fb1, st + NST - 1, st + 1);
      wait_vm<CH * (NST - 3)>();
      __builtin_amdgcn_s_barrier();
      body(fb1, fb0, st + NST, st + 2);
    }
    for (; st < nstage; st += 2) {
      wait_for(st + 1);
      __builtin_amdgcn_s_barrier();
      body(fb0, fb1, st + NST - 1 < nstage ? st + NST - 1 : -1, st + 1 < nstage ? st + 1 : -1);
      if (st + 1 < nstage) {
        wait_for(st + 2);
        __builtin_amdgcn_s_barrier();
        body(fb1, fb0, st + NST < nstage ? st + NST : -1, st + 2 < nstage ? st + 2 : -1);
      }
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    const int otm = tm, otn = tn, osplit = split;
    ++round;
    const int next = round * G + lslot;
    if (next < items) {
      decode(next);
      setup();
      issue_first();
    }
    TO* C = static_cast<TO*>(g.C) + static_cast<long long>(osplit) * g.sCz;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int m = otm * BM + wm0 + 16 * i + r;
      if (m >= g.M) continue;
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        const int n0 = otn * BN + wn0 + 16 * j + 4 * grp;
        if (n0 >= g.N) continue;
        long long col = n0;
        if constexpr (COLBLK) {
          const int blk = n0 / g.nblk;
          col += static_cast<long long>(blk) * (g.sCnb - g.nblk);
        }
        TO* p = C + static_cast<long long>(m) * g.ldcm + col;
        if constexpr (std::is_same<TO, double>::value) {
          if (n0 + 3 < g.N && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
            *reinterpret_cast<double2*>(p) = double2{acc[i][j][0], acc[i][j][1]};
            *reinterpret_cast<double2*>(p + 2) = double2{acc[i][j][2], acc[i][j][3]};
            continue;
          }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (n0 + e < g.N) store_out<TO>(p + e, static_cast<double>(acc[i][j][e]));
      }
    }
    if (next >= items) break;
    item = next;
  }
}

}  // namespace rr
