// librerdmft_gpu, wide direct-to-LDS bf16 GEMM tiles. These kernels keep more accumulators per wave
// (64 x 208 or 128 x 112 per wave, 208-224 accumulator registers) than the architectural VGPR file
// holds next to the fragments and addresses, so this translation unit is compiled with the
// accumulators in the AGPR form (the rest of the library keeps them in VGPRs; see build.py).
#include "rr_gemm.hpp"

namespace rr {

template <typename TO>
static bool launch_wide(int v, const GemmArgs& g, int grid, hipStream_t s) {
  switch (v) {
    case 17: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 128, 416, 2, 3>), dim3(grid), dim3(256), 0, s, g); return true;
    case 18: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 256, 224, 2, 3>), dim3(grid), dim3(256), 0, s, g); return true;
    case 19: hipLaunchKernelGGL((gemm_nt_dma_kernel<TO, 128, 208, 4, 3>), dim3(grid), dim3(256), 0, s, g); return true;
    case 23: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 128, 208, 4, 3>), dim3(grid), dim3(256), 0, s, g); return true;
    case 24: hipLaunchKernelGGL((gemm_nt_dma2_kernel<TO, 128, 416, 2, 3>), dim3(grid), dim3(256), 0, s, g); return true;
    default: return false;
  }
}

// bf16 operands; out_f64 selects the fp64 output type (else bf16). Returns false for an unknown v.
bool launch_dma_agpr(int v, bool out_f64, const GemmArgs& g, int grid, hipStream_t s) {
  return out_f64 ? launch_wide<double>(v, g, grid, s) : launch_wide<bf16_t>(v, g, grid, s);
}

}  // namespace rr
