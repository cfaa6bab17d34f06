// Validation build (compiled with -DRR_VALIDATE only; empty otherwise).
//
// In this mode the library is linked against a host-memory stand-in for the HIP runtime and the
// two solver libraries (tests/validate/hip_mock.cpp) and no kernel is launched: every
// hipLaunchKernelGGL site instead checks that each pointer argument is null or lies inside an
// allocation that is still live, and every GEMM launch checks the highest element its operands
// and outputs can touch (from M, N, K, the leading dimensions, batch / split / k-batch strides and
// the second-output, column-block and row-block mappings) against the size of the allocation the
// pointer lies in. A whole sequence of entry-point calls can then be replayed on the CPU at full
// size; the runtime stand-in reports every violation with the launch site.
//
// Values that the host reads back from device memory to steer control flow (norms, traces,
// symmetry flags, solver status) pass through rr_control(): the stand-in may substitute them so
// the replay takes the branches a real run takes.
#pragma once

#ifdef RR_VALIDATE

#include <cstddef>
#include <type_traits>

extern "C" {
// p is null or lies inside a live allocation; otherwise a violation is recorded.
void rrv_check_live(const void* p, const char* what, const char* where);
// [p, p + bytes) lies inside one live allocation; otherwise a violation is recorded.
void rrv_check_range(const void* p, unsigned long long bytes, const char* what, const char* where);
// A value read back from the device for control flow; the stand-in may substitute it.
double rrv_control(const char* tag, double device_value);
}

namespace rrv {

template <typename T>
inline void arg(const char* where, T* p) {
  rrv_check_live(p, "kernel pointer argument", where);
}
template <typename T>
inline void arg(const char*, const T&) {}

template <typename G, typename B, typename S, typename Q, typename... A>
inline void launch(const char* where, const G&, const B&, const S&, const Q&, const A&... a) {
  (arg(where, a), ...);
}

}  // namespace rrv

#define RRV_STR2(x) #x
#define RRV_STR(x) RRV_STR2(x)
#undef hipLaunchKernelGGL
#define hipLaunchKernelGGL(K, ...) ::rrv::launch(__FILE__ ":" RRV_STR(__LINE__) " " #K, __VA_ARGS__)

#define RR_CONTROL(tag, var) \
  ((var) = static_cast<std::remove_reference_t<decltype(var)>>(rrv_control((tag), static_cast<double>(var))))
// A caller's device array of the extent the entry point's contract gives it.
#define RR_VALIDATE_EXTENT(p, bytes, what) rrv_check_range((p), (unsigned long long)(bytes), (what), __FILE__)

#else

#define RR_CONTROL(tag, var) ((void)0)
#define RR_VALIDATE_EXTENT(p, bytes, what) ((void)0)

#endif
