// Drop-in check of the C++ adapter (rerdmft_amd/host/rerdmft_gpu.hpp): the same templated RHF SCF
// loop (the structure of NonRelHartreeFock.cpp:118-198) is run twice, once with the GPU ERI source
// and once with the CPU oracle's nonRelFockMatrix, and the converged energies must agree to 1e-10.
// Built by rerdmft_amd/build.py; run by tests/test_gpu_cpp_adapter.py.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../../rerdmft_amd/host/rerdmft_gpu.hpp"

// Minimal row-major matrix with the interface of the reference's Matrix<T>.
template <typename T>
class Matrix {
 public:
  Matrix() = default;
  Matrix(std::size_t r, std::size_t c, T fill = T{}) : r_(r), c_(c), d_(r * c, fill) {}
  std::size_t rows() const { return r_; }
  std::size_t cols() const { return c_; }
  T& operator()(std::size_t i, std::size_t j) { return d_[i * c_ + j]; }
  const T& operator()(std::size_t i, std::size_t j) const { return d_[i * c_ + j]; }
  T* data() { return d_.data(); }
  const T* data() const { return d_.data(); }

 private:
  std::size_t r_ = 0, c_ = 0;
  std::vector<T> d_;
};
using M = Matrix<double>;

M mul(const M& a, const M& b, bool ta = false) {
  const std::size_t m = ta ? a.cols() : a.rows(), k = ta ? a.rows() : a.cols();
  M c(m, b.cols());
  for (std::size_t i = 0; i < m; ++i)
    for (std::size_t p = 0; p < k; ++p) {
      const double av = ta ? a(p, i) : a(i, p);
      for (std::size_t j = 0; j < b.cols(); ++j) c(i, j) += av * b(p, j);
    }
  return c;
}

// Cyclic Jacobi: eigenvalues ascending, eigenvectors in columns.
void eigh(const M& a_in, std::vector<double>& w, M& v) {
  const std::size_t n = a_in.rows();
  M a = a_in;
  v = M(n, n);
  for (std::size_t i = 0; i < n; ++i) v(i, i) = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (std::size_t p = 0; p < n; ++p)
      for (std::size_t q = p + 1; q < n; ++q) off += a(p, q) * a(p, q);
    if (off < 1e-30) break;
    for (std::size_t p = 0; p < n; ++p)
      for (std::size_t q = p + 1; q < n; ++q) {
        if (std::fabs(a(p, q)) < 1e-300) continue;
        const double th = 0.5 * std::atan2(2 * a(p, q), a(q, q) - a(p, p));
        const double c = std::cos(th), s = std::sin(th);
        for (std::size_t k = 0; k < n; ++k) {
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (std::size_t k = 0; k < n; ++k) {
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (std::size_t k = 0; k < n; ++k) {
          const double vkp = v(k, p), vkq = v(k, q);
          v(k, p) = c * vkp - s * vkq;
          v(k, q) = s * vkp + c * vkq;
        }
      }
  }
  std::vector<std::size_t> idx(n);
  for (std::size_t i = 0; i < n; ++i) idx[i] = i;
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t j = i + 1; j < n; ++j)
      if (a(idx[j], idx[j]) < a(idx[i], idx[i])) std::swap(idx[i], idx[j]);
  M vs(n, n);
  w.resize(n);
  for (std::size_t j = 0; j < n; ++j) {
    w[j] = a(idx[j], idx[j]);
    for (std::size_t i = 0; i < n; ++i) vs(i, j) = v(i, idx[j]);
  }
  v = vs;
}

struct Lcg {
  std::uint64_t s;
  double next() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return static_cast<double>(s >> 11) / 9007199254740992.0 - 0.5;
  }
};

// CPU oracle source: same overload name, resolved for a different source type.
using OracleFn = int (*)(std::size_t, std::size_t, const double*, const double*, std::size_t, const double*, const double*,
                         const double*, double*, double*, double*);
struct OracleSource {
  OracleFn fn;
  std::size_t nl, naux;
  const M* eri3;
  M x, s;
};
M nonRelFockMatrix(const M& h, const OracleSource& src, const M& p) {
  M f(h.rows(), h.cols());
  if (src.fn(src.nl, src.naux, src.eri3->data(), src.x.data(), src.x.cols(), src.s.data(), h.data(), p.data(), f.data(),
             nullptr, nullptr) != 0)
    throw std::runtime_error("oracle failed");
  return f;
}

// Templated RHF SCF loop (NonRelHartreeFock.cpp:118-198 structure, without DIIS): the only
// dependency on the two-electron representation is the nonRelFockMatrix overload.
template <class Eri>
double scf(const M& h, const Eri& eri, const M& x, std::size_t nocc, int* iters) {
  const std::size_t n = h.rows();
  auto density = [&](const M& f) {
    std::vector<double> w;
    M y;
    eigh(mul(x, mul(f, x), true), w, y);
    M c = mul(x, y);
    M p(n, n);
    for (std::size_t a = 0; a < n; ++a)
      for (std::size_t b = 0; b < n; ++b)
        for (std::size_t i = 0; i < nocc; ++i) p(a, b) += 2.0 * c(a, i) * c(b, i);
    return p;
  };
  M p = density(h);  // core-Hamiltonian guess
  double e_old = 0.0, e = 0.0;
  for (int it = 0; it < 200; ++it) {
    M f = nonRelFockMatrix(h, eri, p);
    e = 0.0;
    for (std::size_t a = 0; a < n; ++a)
      for (std::size_t b = 0; b < n; ++b) e += 0.5 * p(a, b) * (h(a, b) + f(a, b));
    *iters = it + 1;
    if (it > 0 && std::fabs(e - e_old) < 1e-12) break;
    e_old = e;
    M p_new = density(f);
    for (std::size_t i = 0; i < n * n; ++i) p.data()[i] = 0.5 * p.data()[i] + 0.5 * p_new.data()[i];
  }
  return e;
}

// The same loop with DIIS on the Fock matrix (NonRelHartreeFock.cpp:118-198 with its default DIIS):
// the commutator error and the extrapolation run through the GPU counterparts of the reference's
// scfCommutatorError / Diis<double>.
template <class Eri>
double scf_diis(const M& h, const Eri& eri, const M& x, const M& s, std::size_t nocc,
                const std::shared_ptr<rerdmft::gpu::Device>& dev, int* iters) {
  const std::size_t n = h.rows();
  auto density = [&](const M& f) {
    std::vector<double> w;
    M y;
    eigh(mul(x, mul(f, x), true), w, y);
    M c = mul(x, y);
    M p(n, n);
    for (std::size_t a = 0; a < n; ++a)
      for (std::size_t b = 0; b < n; ++b)
        for (std::size_t i = 0; i < nocc; ++i) p(a, b) += 2.0 * c(a, i) * c(b, i);
    return p;
  };
  rerdmft::gpu::GpuDiis<M> diis(dev, 5);
  M p = density(h);
  double e_old = 0.0, e = 0.0;
  for (int it = 0; it < 200; ++it) {
    M f = nonRelFockMatrix(h, eri, p);
    e = 0.0;
    for (std::size_t a = 0; a < n; ++a)
      for (std::size_t b = 0; b < n; ++b) e += 0.5 * p(a, b) * (h(a, b) + f(a, b));
    *iters = it + 1;
    if (it > 0 && std::fabs(e - e_old) < 1e-12) break;
    e_old = e;
    const M err = rerdmft::gpu::scfCommutatorError(*dev, f, p, s);
    p = density(diis.extrapolate(err, f));
  }
  return e;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <liboracle.so>\n", argv[0]);
    return 2;
  }
  void* lib = dlopen(argv[1], RTLD_NOW);
  if (!lib) {
    std::fprintf(stderr, "cannot load oracle: %s\n", dlerror());
    return 2;
  }
  const std::size_t n = 30, naux = 90, nocc = 5;
  Lcg rng{20261203ULL};
  M b(naux, n * n);
  for (std::size_t P = 0; P < naux; ++P)
    for (std::size_t i = 0; i < n; ++i)
      for (std::size_t j = i; j < n; ++j) {
        const double v = 0.2 * rng.next() * std::exp(-std::fabs(double(i) - double(j)) / 6.0);
        b(P, i * n + j) = v;
        b(P, j * n + i) = v;
      }
  M h(n, n), s(n, n);
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t j = i; j < n; ++j) {
      h(i, j) = h(j, i) = (i == j ? -2.0 + 0.1 * double(i) : 0.3 * rng.next());
      s(i, j) = s(j, i) = (i == j ? 1.0 : 0.1 * rng.next() * std::exp(-std::fabs(double(i) - double(j))));
    }
  std::vector<double> w;
  M u;
  eigh(s, w, u);
  M x(n, n);  // S^-1/2
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t j = 0; j < n; ++j)
      for (std::size_t k = 0; k < n; ++k) x(i, j) += u(i, k) * u(j, k) / std::sqrt(w[k]);

  auto dev = std::make_shared<rerdmft::gpu::Device>(0);
  rerdmft::gpu::GpuRiNonRelEriSource<M> gsrc(dev, n, b, x, s, RR_FP64);
  OracleSource osrc{reinterpret_cast<OracleFn>(dlsym(lib, "oracle_nonrel_fock_ri")), n, naux, &b, x, s};
  if (!osrc.fn) return 2;
  int it_gpu = 0, it_cpu = 0;
  const double e_gpu = scf(h, gsrc, x, nocc, &it_gpu);
  const double e_cpu = scf(h, osrc, x, nocc, &it_cpu);
  std::printf("E_gpu = %.12f (%d iterations)\nE_cpu = %.12f (%d iterations)\n|dE| = %.3e\n", e_gpu, it_gpu, e_cpu, it_cpu,
              std::fabs(e_gpu - e_cpu));
  bool ok = std::fabs(e_gpu - e_cpu) < 1e-10;
  // DIIS through the GPU counterparts, on a weaker two-electron part where the SCF converges: the
  // energy must equal the (slower) linearly mixed loop's
  M b2 = b;
  for (std::size_t i = 0; i < naux * n * n; ++i) b2.data()[i] *= 0.25;
  rerdmft::gpu::GpuRiNonRelEriSource<M> gsrc2(dev, n, b2, x, s, RR_FP64);
  int it_mix = 0, it_diis = 0;
  const double e_mix = scf(h, gsrc2, x, nocc, &it_mix);
  const double e_diis = scf_diis(h, gsrc2, x, s, nocc, dev, &it_diis);
  std::printf("weaker system: E_mix = %.12f (%d iterations), E_diis = %.12f (%d iterations)\n", e_mix, it_mix, e_diis,
              it_diis);
  ok = ok && it_mix < 200 && it_diis < it_mix && std::fabs(e_diis - e_mix) < 1e-9;
  // shape error propagates as the reference's message
  try {
    M wrong(n + 1, n + 1);
    (void)nonRelFockMatrix(wrong, gsrc, wrong);
    ok = false;
  } catch (const std::runtime_error& e) {
    std::printf("shape error: %s\n", e.what());
  }
  std::printf(ok ? "[PASS]\n" : "[FAIL]\n");
  return ok ? 0 : 1;
}
