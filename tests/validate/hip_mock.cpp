// Host-memory stand-in for the HIP runtime, rocBLAS and rocSOLVER, for the validation build of
// librerdmft_gpu (-DRR_VALIDATE, see rerdmft_amd/csrc/rr_validate.hpp). TEST INFRASTRUCTURE ONLY.
//
// No device code runs. "Device" allocations are anonymous host mappings that are never touched
// unless a copy or fill of at most kCopyLimit bytes writes them, and that are never returned to the
// system: a freed allocation stays mapped without access and stays in the registry as freed, so a
// pointer into it is recognised at every later use. The library's launch sites report each pointer
// argument (rrv_check_live) and each GEMM operand / output extent (rrv_check_range); copies, fills
// and the solver calls are checked here. Pointers recorded while a stream is being captured are
// checked again whenever the instantiated graph is launched. Every violation is printed with its
// site and counted (rrv_violations).
//
// Control values (rrv_control): the replay sets RRV_CONTROL="tag=value,..." to substitute the values
// the host reads back from the device (norms, traces, symmetry flags). The eigensolvers write
// RRV_EIG_TOP eigenvalues equal to 1 at the top of an ascending spectrum of zeros.
#include <hip/hip_runtime_api.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <sys/mman.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace {

constexpr size_t kCopyLimit = size_t(64) << 20;

struct Alloc {
  size_t size = 0;
  bool live = true;
  long serial = 0;
};

struct State {
  std::mutex mu;
  std::map<uintptr_t, Alloc> allocs;  // base -> allocation
  long serial = 0;
  long violations = 0;
  long launches_checked = 0;
  bool capturing = false;
  std::vector<std::pair<const void*, unsigned long long>> capture;  // (pointer, bytes; 0 = liveness only)
  std::string last_violation;
};

State& st() {
  static State s;
  return s;
}

struct Graph {
  std::vector<std::pair<const void*, unsigned long long>> uses;
};

// Allocation containing p (p may equal one past the end); null if none.
const std::pair<const uintptr_t, Alloc>* find(const void* p) {
  auto& m = st().allocs;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = m.upper_bound(a);
  if (it == m.begin()) return nullptr;
  --it;
  if (a >= it->first && a <= it->first + it->second.size) return &*it;
  return nullptr;
}

void violation(const std::string& msg) {
  State& s = st();
  ++s.violations;
  s.last_violation = msg;
  std::fprintf(stderr, "RRV VIOLATION: %s\n", msg.c_str());
  std::fflush(stderr);
}

std::string hexp(const void* p) {
  char b[32];
  std::snprintf(b, sizeof(b), "%p", p);
  return b;
}

// Checks [p, p + bytes) (bytes = 0: liveness only). Unknown pointers are violations when `device`
// (a kernel argument must be device memory); for copies they are host memory and pass.
bool check(const void* p, unsigned long long bytes, const char* what, const char* where, bool device) {
  if (!p) {
    if (bytes == 0) return true;
    violation(std::string(what) + ": null pointer for " + std::to_string(bytes) + " bytes at " + where);
    return false;
  }
  const auto* a = find(p);
  if (!a) {
    if (device) violation(std::string(what) + ": " + hexp(p) + " is not device memory at " + where);
    return !device;
  }
  if (!a->second.live) {
    violation(std::string(what) + ": " + hexp(p) + " points into a freed allocation (#" +
              std::to_string(a->second.serial) + ", " + std::to_string(a->second.size) + " bytes) at " + where);
    return false;
  }
  const uintptr_t off = reinterpret_cast<uintptr_t>(p) - a->first;
  if (off + bytes > a->second.size) {
    violation(std::string(what) + ": [" + std::to_string(off) + ", " + std::to_string(off + bytes) +
              ") exceeds the allocation (#" + std::to_string(a->second.serial) + ", " +
              std::to_string(a->second.size) + " bytes) by " + std::to_string(off + bytes - a->second.size) +
              " bytes at " + where);
    return false;
  }
  return true;
}

void note_capture(const void* p, unsigned long long bytes) {
  State& s = st();
  if (s.capturing && p) s.capture.push_back({p, bytes});
}

double control_value(const char* tag, double dflt) {
  const char* e = std::getenv("RRV_CONTROL");
  if (!e) return dflt;
  const std::string all(e), key = std::string(tag) + "=";
  size_t pos = 0;
  while (pos < all.size()) {
    size_t end = all.find(',', pos);
    if (end == std::string::npos) end = all.size();
    const std::string item = all.substr(pos, end - pos);
    if (item.compare(0, key.size(), key) == 0) return std::atof(item.c_str() + key.size());
    pos = end + 1;
  }
  return dflt;
}

bool device_ptr(const void* p) {
  const auto* a = find(p);
  return a && a->second.live;
}

void copy_checked(void* dst, const void* src, size_t n, const char* where) {
  std::lock_guard<std::mutex> g(st().mu);
  const bool dd = find(dst) != nullptr, sd = find(src) != nullptr;
  bool ok = check(dst, n, "copy destination", where, false) & check(src, n, "copy source", where, false);
  note_capture(dd ? dst : nullptr, n);
  note_capture(sd ? src : nullptr, n);
  if (ok && n && n <= kCopyLimit && dst && src) std::memmove(dst, src, n);
}

}  // namespace

extern "C" {

// ---- validation interface (rr_validate.hpp) ------------------------------------------------------
void rrv_check_live(const void* p, const char* what, const char* where) {
  std::lock_guard<std::mutex> g(st().mu);
  ++st().launches_checked;
  check(p, 0, what, where, true);
  note_capture(p, 0);
}
void rrv_check_range(const void* p, unsigned long long bytes, const char* what, const char* where) {
  std::lock_guard<std::mutex> g(st().mu);
  check(p, bytes, what, where, true);
  note_capture(p, bytes);
}
double rrv_control(const char* tag, double device_value) { return control_value(tag, device_value); }
long rrv_violations() { return st().violations; }
long rrv_checks() { return st().launches_checked; }
const char* rrv_last_violation() { return st().last_violation.c_str(); }
// Registers caller-side "device" memory (the replay allocates through hipMalloc, so this is only
// for completeness) and returns the number of live allocations.
long rrv_live_allocations() {
  std::lock_guard<std::mutex> g(st().mu);
  long n = 0;
  for (auto& kv : st().allocs) n += kv.second.live ? 1 : 0;
  return n;
}

// ---- HIP runtime -----------------------------------------------------------------------------------
void** __hipRegisterFatBinary(const void*) {
  static void* handle[4] = {nullptr, nullptr, nullptr, nullptr};
  return handle;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterVar(void**, void*, char*, const char*, int, size_t, int, int) {}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, void*, void*, void*, void*, int*) {}
hipError_t __hipPushCallConfiguration(dim3, dim3, size_t, hipStream_t) { return hipSuccess; }
hipError_t __hipPopCallConfiguration(dim3*, dim3*, size_t*, hipStream_t*) { return hipSuccess; }

hipError_t hipLaunchKernel(const void*, dim3, dim3, void**, size_t, hipStream_t) {
  violation("a kernel was launched past the validation hooks (hipLaunchKernel)");
  return hipSuccess;
}

hipError_t hipMalloc(void** p, size_t size) {
  if (!p) return hipErrorInvalidValue;
  const size_t len = ((size ? size : 1) + 4095) / 4096 * 4096;
  void* m = mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (m == MAP_FAILED) return hipErrorOutOfMemory;
  std::lock_guard<std::mutex> g(st().mu);
  Alloc a;
  a.size = size;
  a.serial = ++st().serial;
  st().allocs[reinterpret_cast<uintptr_t>(m)] = a;
  *p = m;
  return hipSuccess;
}

// Device variables: each symbol gets one allocation of kSymbolBytes (the stand-in does not know
// the variable's size; the library's largest is its 160 MiB scratch region).
hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
  constexpr size_t kSymbolBytes = size_t(256) << 20;
  static std::map<const void*, void*> symbols;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  auto it = symbols.find(symbol);
  if (it == symbols.end()) {
    void* m = nullptr;
    const hipError_t e = hipMalloc(&m, kSymbolBytes);
    if (e != hipSuccess) return e;
    it = symbols.emplace(symbol, m).first;
  }
  if (p) *p = it->second;
  return hipSuccess;
}

hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> g(st().mu);
  auto it = st().allocs.find(reinterpret_cast<uintptr_t>(p));
  if (it == st().allocs.end()) {
    violation("hipFree of a pointer that is not the base of an allocation: " + hexp(p));
    return hipErrorInvalidValue;
  }
  if (!it->second.live) {
    violation("double hipFree of allocation #" + std::to_string(it->second.serial));
    return hipErrorInvalidValue;
  }
  it->second.live = false;
  const size_t len = ((it->second.size ? it->second.size : 1) + 4095) / 4096 * 4096;
  madvise(p, len, MADV_DONTNEED);
  mprotect(p, len, PROT_NONE);
  return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind) {
  copy_checked(dst, src, n, "hipMemcpy");
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
  copy_checked(dst, src, n, "hipMemcpyAsync");
  return hipSuccess;
}
hipError_t hipMemcpy2DAsync(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind, hipStream_t) {
  if (height == 0 || width == 0) return hipSuccess;
  std::lock_guard<std::mutex> g(st().mu);
  const bool dd = find(dst) != nullptr, sd = find(src) != nullptr;
  const unsigned long long db = (height - 1) * dpitch + width, sb = (height - 1) * spitch + width;
  const bool ok = check(dst, db, "2D copy destination", "hipMemcpy2DAsync", false) &
                  check(src, sb, "2D copy source", "hipMemcpy2DAsync", false);
  note_capture(dd ? dst : nullptr, db);
  note_capture(sd ? src : nullptr, sb);
  if (ok && height * width <= kCopyLimit)
    for (size_t r = 0; r < height; ++r)
      std::memmove(static_cast<char*>(dst) + r * dpitch, static_cast<const char*>(src) + r * spitch, width);
  return hipSuccess;
}
hipError_t hipMemset(void* dst, int v, size_t n) {
  std::lock_guard<std::mutex> g(st().mu);
  if (check(dst, n, "fill destination", "hipMemset", true) && n <= kCopyLimit) std::memset(dst, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int v, size_t n, hipStream_t) {
  std::lock_guard<std::mutex> g(st().mu);
  note_capture(dst, n);
  if (check(dst, n, "fill destination", "hipMemsetAsync", true) && n <= kCopyLimit) std::memset(dst, v, n);
  return hipSuccess;
}

hipError_t hipPointerGetAttributes(hipPointerAttribute_t* attr, const void* p) {
  std::lock_guard<std::mutex> g(st().mu);
  if (!attr) return hipErrorInvalidValue;
  std::memset(attr, 0, sizeof(*attr));
  if (!device_ptr(p)) return hipErrorInvalidValue;
  attr->type = hipMemoryTypeDevice;
  attr->device = 0;
  attr->devicePointer = const_cast<void*>(p);
  return hipSuccess;
}

hipError_t hipGetDeviceCount(int* n) {
  if (n) *n = 1;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t attr, int) {
  if (!v) return hipErrorInvalidValue;
  *v = attr == hipDeviceAttributeMultiprocessorCount ? 256 : 0;  // an MI355X has 256 compute units
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "hip mock error"; }

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  static char token[64];
  static int next = 0;
  if (s) *s = reinterpret_cast<hipStream_t>(token + (next++ % 64));
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }

hipError_t hipEventCreate(hipEvent_t* e) {
  if (e) *e = reinterpret_cast<hipEvent_t>(new char[1]);
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete[] reinterpret_cast<char*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  if (ms) *ms = 1.0f;
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  std::lock_guard<std::mutex> g(st().mu);
  st().capturing = true;
  st().capture.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* graph) {
  std::lock_guard<std::mutex> g(st().mu);
  Graph* gr = new Graph;
  gr->uses.swap(st().capture);
  st().capturing = false;
  if (graph) *graph = reinterpret_cast<hipGraph_t>(gr);
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t graph, hipGraphNode_t*, char*, size_t) {
  Graph* e = new Graph(*reinterpret_cast<Graph*>(graph));
  if (exec) *exec = reinterpret_cast<hipGraphExec_t>(e);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
  std::lock_guard<std::mutex> g(st().mu);
  for (const auto& u : reinterpret_cast<Graph*>(exec)->uses)
    check(u.first, u.second, "graph replay", "hipGraphLaunch", true);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t graph) {
  delete reinterpret_cast<Graph*>(graph);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
  delete reinterpret_cast<Graph*>(exec);
  return hipSuccess;
}

// ---- rocBLAS / rocSOLVER (column-major operands) ------------------------------------------------
rocblas_status rocblas_create_handle(rocblas_handle* h) {
  if (h) *h = reinterpret_cast<rocblas_handle>(new char[1]);
  return rocblas_status_success;
}
rocblas_status rocblas_destroy_handle(rocblas_handle h) {
  delete[] reinterpret_cast<char*>(h);
  return rocblas_status_success;
}
rocblas_status rocblas_set_stream(rocblas_handle, hipStream_t) { return rocblas_status_success; }

namespace {
// rows x cols column-major array with leading dimension ld, element size es.
bool check_cm(const void* p, long rows, long cols, long ld, size_t es, const char* what, const char* where) {
  if (rows <= 0 || cols <= 0) return true;
  if (ld < rows) {
    violation(std::string(what) + ": leading dimension " + std::to_string(ld) + " below " + std::to_string(rows) +
              " rows at " + where);
    return false;
  }
  std::lock_guard<std::mutex> g(st().mu);
  const unsigned long long bytes = ((unsigned long long)(cols - 1) * ld + rows) * es;
  note_capture(p, bytes);
  return check(p, bytes, what, where, true);
}
void gemm_check(rocblas_operation ta, rocblas_operation tb, long m, long n, long k, const void* A, long lda,
                const void* B, long ldb, const void* C, long ldc, size_t es, const char* where) {
  if (k > 0) {
    check_cm(A, ta == rocblas_operation_none ? m : k, ta == rocblas_operation_none ? k : m, lda, es, "gemm A", where);
    check_cm(B, tb == rocblas_operation_none ? k : n, tb == rocblas_operation_none ? n : k, ldb, es, "gemm B", where);
  }
  check_cm(C, m, n, ldc, es, "gemm C", where);
}
void eig_fill(long n, double* w) {
  long top = (long)control_value("eig_top", -1.0);
  if (top < 0) {
    const char* e = std::getenv("RRV_EIG_TOP");
    top = e ? std::atol(e) : 0;
  }
  if (!w || n <= 0) return;
  for (long i = 0; i < n; ++i) w[i] = (i >= n - top) ? 1.0 : 0.0;
}
}  // namespace

rocblas_status rocblas_zgemm(rocblas_handle, rocblas_operation ta, rocblas_operation tb, rocblas_int m, rocblas_int n,
                             rocblas_int k, const rocblas_double_complex*, const rocblas_double_complex* A,
                             rocblas_int lda, const rocblas_double_complex* B, rocblas_int ldb,
                             const rocblas_double_complex*, rocblas_double_complex* C, rocblas_int ldc) {
  gemm_check(ta, tb, m, n, k, A, lda, B, ldb, C, ldc, 16, "rocblas_zgemm");
  return rocblas_status_success;
}
rocblas_status rocblas_gemm_ex(rocblas_handle, rocblas_operation ta, rocblas_operation tb, rocblas_int m, rocblas_int n,
                               rocblas_int k, const void*, const void* A, rocblas_datatype, rocblas_int lda, const void* B,
                               rocblas_datatype, rocblas_int ldb, const void*, const void* C, rocblas_datatype,
                               rocblas_int ldc, void* D, rocblas_datatype, rocblas_int ldd, rocblas_datatype,
                               rocblas_gemm_algo, int32_t, uint32_t) {
  // f64 operands only (the library's sole use)
  gemm_check(ta, tb, m, n, k, A, lda, B, ldb, C, ldc, 8, "rocblas_gemm_ex C");
  gemm_check(ta, tb, m, n, 0, A, lda, B, ldb, D, ldd, 8, "rocblas_gemm_ex D");
  return rocblas_status_success;
}
rocblas_status rocblas_ztrsm(rocblas_handle, rocblas_side side, rocblas_fill, rocblas_operation, rocblas_diagonal,
                             rocblas_int m, rocblas_int n, const rocblas_double_complex*,
                             const rocblas_double_complex* A, rocblas_int lda, rocblas_double_complex* B,
                             rocblas_int ldb) {
  const long ka = side == rocblas_side_left ? m : n;
  check_cm(A, ka, ka, lda, 16, "trsm A", "rocblas_ztrsm");
  check_cm(B, m, n, ldb, 16, "trsm B", "rocblas_ztrsm");
  return rocblas_status_success;
}
rocblas_status rocblas_dtrsm(rocblas_handle, rocblas_side side, rocblas_fill, rocblas_operation, rocblas_diagonal,
                             rocblas_int m, rocblas_int n, const double*, const double* A, rocblas_int lda, double* B,
                             rocblas_int ldb) {
  const long ka = side == rocblas_side_left ? m : n;
  check_cm(A, ka, ka, lda, 8, "trsm A", "rocblas_dtrsm");
  check_cm(B, m, n, ldb, 8, "trsm B", "rocblas_dtrsm");
  return rocblas_status_success;
}
rocblas_status rocsolver_dpotrf(rocblas_handle, const rocblas_fill, const rocblas_int n, double* A, const rocblas_int lda,
                                rocblas_int* info) {
  check_cm(A, n, n, lda, 8, "potrf A", "rocsolver_dpotrf");
  check_cm(info, 1, 1, 1, sizeof(rocblas_int), "potrf info", "rocsolver_dpotrf");
  return rocblas_status_success;
}
rocblas_status rocsolver_zpotrf(rocblas_handle, const rocblas_fill, const rocblas_int n, rocblas_double_complex* A,
                                const rocblas_int lda, rocblas_int* info) {
  check_cm(A, n, n, lda, 16, "potrf A", "rocsolver_zpotrf");
  check_cm(info, 1, 1, 1, sizeof(rocblas_int), "potrf info", "rocsolver_zpotrf");
  return rocblas_status_success;
}
rocblas_status rocsolver_dsyevd(rocblas_handle, const rocblas_evect, const rocblas_fill, const rocblas_int n, double* A,
                                const rocblas_int lda, double* D, double* E, rocblas_int* info) {
  check_cm(A, n, n, lda, 8, "syevd A", "rocsolver_dsyevd");
  if (check_cm(D, n, 1, n, 8, "syevd D", "rocsolver_dsyevd")) eig_fill(n, D);
  check_cm(E, n, 1, n, 8, "syevd E", "rocsolver_dsyevd");
  check_cm(info, 1, 1, 1, sizeof(rocblas_int), "syevd info", "rocsolver_dsyevd");
  return rocblas_status_success;
}
rocblas_status rocsolver_zheevd(rocblas_handle, const rocblas_evect, const rocblas_fill, const rocblas_int n,
                                rocblas_double_complex* A, const rocblas_int lda, double* D, double* E,
                                rocblas_int* info) {
  check_cm(A, n, n, lda, 16, "heevd A", "rocsolver_zheevd");
  if (check_cm(D, n, 1, n, 8, "heevd D", "rocsolver_zheevd")) eig_fill(n, D);
  check_cm(E, n, 1, n, 8, "heevd E", "rocsolver_zheevd");
  check_cm(info, 1, 1, 1, sizeof(rocblas_int), "heevd info", "rocsolver_zheevd");
  return rocblas_status_success;
}

}  // extern "C"
