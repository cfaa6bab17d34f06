// CPU oracle, functional stage: the RI generalized-Fock GEMM and the RI sigma-vector engine.
// TEST INFRASTRUCTURE ONLY (see rr_oracle.cpp). Plain loops over FP64 / complex<double>, OpenMP over
// output rows; each function names the reference routine it restates.
//
// Conventions (src/Hessian_opt/RiSigmaGemm.h:13-22): B(P,p,q) at b[P*n*n + p*n + q];
// eri(a,b,c,d) = sum_P B(P,a,c) B(P,b,d), unconjugated.
#include <omp.h>

#include <complex>
#include <cstddef>
#include <cstdlib>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using C = std::complex<double>;
thread_local std::string g_err_sigma;

// Mutation switch for the oracle's own tests (RR_ORACLE_PLANT=<name>); unset in every normal run.
bool plantedSigma(const char* name) {
  const char* e = std::getenv("RR_ORACLE_PLANT");
  return e && std::string(e) == name;
}

inline double conjv(double x) { return x; }
inline C conjv(C x) { return std::conj(x); }
inline double re(double x) { return x; }
inline double re(C x) { return x.real(); }

template <typename T>
const T* as(const double* p) {
  return reinterpret_cast<const T*>(p);
}
template <typename T>
T* as(double* p) {
  return reinterpret_cast<T*>(p);
}

// ---- primitives (src/Hessian_opt/RiSigmaGemm.cpp) ------------------------------------------------
template <typename T>
struct Op {
  const T* s = nullptr;
  std::vector<std::uint32_t> idx;
  std::vector<double> w;
  const T* row = nullptr;
  std::vector<double> post;  // n*n, empty: none
};

// value(P,t,x) of an operand (fillOp, RiSigmaGemm.cpp:46-72), including the post factor
template <typename T>
std::vector<T> opValues(const Op<T>& op, std::size_t n, std::size_t naux) {
  const std::size_t n2 = n * n;
  std::vector<T> v(naux * n2);
#pragma omp parallel for collapse(2)
  for (std::size_t P = 0; P < naux; ++P)
    for (std::size_t t = 0; t < n; ++t)
      for (std::size_t x = 0; x < n; ++x) {
        T val = op.s ? op.s[P * n2 + op.idx[t * n + x]] : T(1.0);
        if (!op.w.empty()) val *= op.w[t * n + x];
        if (op.row && !plantedSigma("sigma_row")) val *= op.row[P * n + t];
        v[(P * n + t) * n + x] = val;
      }
  if (op.post.empty() || plantedSigma("sigma_post")) return v;
  std::vector<T> out(naux * n2, T{});
#pragma omp parallel for collapse(2)
  for (std::size_t k = 0; k < naux * n; ++k)
    for (std::size_t x = 0; x < n; ++x) {
      T acc{};
      for (std::size_t a = 0; a < n; ++a) acc += v[k * n + a] * op.post[a * n + x];
      out[k * n + x] = acc;
    }
  return out;
}

// pairContract (RiSigmaGemm.cpp:76-101): target(x,y) += sign sum_{P,t} F(P,t,x) G(P,t,y)
template <typename T>
void pairContract(std::vector<T>& target, double sign, std::size_t n, std::size_t naux, const Op<T>& f, const Op<T>& g) {
  const std::vector<T> fv = opValues(f, n, naux), gv = opValues(g, n, naux);
#pragma omp parallel for collapse(2)
  for (std::size_t x = 0; x < n; ++x)
    for (std::size_t y = 0; y < n; ++y) {
      T acc{};
      for (std::size_t k = 0; k < naux * n; ++k) acc += fv[k * n + x] * gv[k * n + y];
      target[x * n + y] += sign * acc;
    }
}

// diagContract (RiSigmaGemm.cpp:103-121)
template <typename T>
void diagContract(std::vector<T>& target, std::size_t n, std::size_t naux, const T* S, const T* Ua, const T* Ub) {
#pragma omp parallel for collapse(2)
  for (std::size_t a = 0; a < n; ++a)
    for (std::size_t b = 0; b < n; ++b) {
      T acc{};
      for (std::size_t P = 0; P < naux; ++P) {
        const T u = (Ua ? Ua[P * n + a] : T{}) + (Ub ? Ub[P * n + b] : T{});
        acc += S[P * n * n + b * n + a] * u;
      }
      target[a * n + b] += acc;
    }
}

// accumU (RiSigmaGemm.cpp:123-132)
template <typename T>
void accumU(std::vector<T>& U, double alpha, std::size_t n, std::size_t naux, const T* D, const double* Wm) {
#pragma omp parallel for collapse(2)
  for (std::size_t P = 0; P < naux; ++P)
    for (std::size_t x = 0; x < n; ++x) {
      T acc{};
      for (std::size_t t = 0; t < n; ++t) acc += D[P * n + t] * Wm[x * n + t];
      U[P * n + x] += alpha * acc;
    }
}

// ---- toolkit (RiSigmaHessianToolkit.cpp:25-56) -------------------------------------------------
template <typename T>
struct Toolkit {
  std::size_t n = 0, naux = 0;
  const T* b = nullptr;
  std::vector<T> kb1, kb2, bdiag, kb1diag, kb2diag;
};

template <typename T>
Toolkit<T> buildToolkit(std::size_t n, std::size_t naux, const T* b, const double* K) {
  Toolkit<T> t;
  t.n = n, t.naux = naux, t.b = b;
  const std::size_t n2 = n * n;
  t.kb1.assign(naux * n2, T{});
  t.kb2.assign(naux * n2, T{});
  t.bdiag.assign(naux * n, T{});
  t.kb1diag.assign(naux * n, T{});
  t.kb2diag.assign(naux * n, T{});
#pragma omp parallel for
  for (std::size_t P = 0; P < naux; ++P) {
    const T* bp = b + P * n2;
    for (std::size_t a = 0; a < n; ++a)
      for (std::size_t c = 0; c < n; ++c) {
        T s1{}, s2{};
        for (std::size_t s = 0; s < n; ++s) {
          s1 += K[a * n + s] * bp[s * n + c];
          s2 += bp[a * n + s] * K[s * n + c];
        }
        t.kb1[P * n2 + a * n + c] = s1;
        t.kb2[P * n2 + a * n + c] = s2;
      }
    for (std::size_t x = 0; x < n; ++x) {
      t.bdiag[P * n + x] = bp[x * n + x];
      t.kb1diag[P * n + x] = t.kb1[P * n2 + x * n + x];
      t.kb2diag[P * n + x] = t.kb2[P * n2 + x * n + x];
    }
  }
  return t;
}

// ---- riHxFockGemm (RiGemmFock.cpp:53-115) ------------------------------------------------------
template <typename T>
void hxFock(std::size_t n, std::size_t naux, const T* b, const T* h, const double* occ, const double* h2,
            const double* x2, const std::int64_t* pair, const double* l1, const double* l2, T* f) {
  const std::size_t n2 = n * n;
  // W(P,q) = sum_s Bdiag(P,s) h2(s,q)
  std::vector<T> w(naux * n, T{});
  for (std::size_t P = 0; P < naux; ++P)
    for (std::size_t q = 0; q < n; ++q)
      for (std::size_t s = 0; s < n; ++s) w[P * n + q] += b[P * n2 + s * n + s] * h2[s * n + q];
#pragma omp parallel for collapse(2)
  for (std::size_t p = 0; p < n; ++p)
    for (std::size_t q = 0; q < n; ++q) {
      T g{};
      for (std::size_t P = 0; P < naux; ++P) {
        const T* bp = b + P * n2;
        g += bp[q * n + p] * w[P * n + q];  // h2 term
        for (std::size_t t = 0; t < n; ++t) {
          T y = -x2[q * n + t] * bp[q * n + t];
          if (pair) {
            const std::size_t st = static_cast<std::size_t>(pair[t]), qb = static_cast<std::size_t>(pair[q]);
            y += (plantedSigma("hx_l_factor") ? 1.0 : 2.0) * (l1[st * n + qb] + l2[st * n + q]) * bp[st * n + qb];
          }
          g += bp[t * n + p] * y;
        }
      }
      f[p * n + q] = occ[q] * h[q * n + p] + g;
    }
}

// ---- sigma drivers --------------------------------------------------------------------------------
using Z = std::size_t;
using IdxFn = std::function<std::size_t(Z, Z)>;
using WFn = std::function<double(Z, Z)>;

template <typename T>
struct Sigma {
  std::size_t n, naux;
  const Toolkit<T>& t;
  const T* h;
  const double* occ;
  const double* hc;
  const double* xc;
  const double* K;
  const T* fock = nullptr;
  const std::int64_t* pair = nullptr;
  const double* l1 = nullptr;
  const double* l2 = nullptr;

  double HC(Z a, Z b) const { return hc[a * n + b]; }
  double XC(Z a, Z b) const { return xc[a * n + b]; }
  double KK(Z a, Z b) const { return K[a * n + b]; }
  double L1(Z a, Z b) const { return l1[a * n + b]; }
  double L2(Z a, Z b) const { return l2[a * n + b]; }
  Z pr(Z a) const { return static_cast<Z>(pair[a]); }

  Op<T> op(const T* s, IdxFn idx, WFn w = nullptr) const {
    Op<T> o;
    o.s = s;
    if (s) {
      o.idx.resize(n * n);
      for (Z a = 0; a < n; ++a)
        for (Z x = 0; x < n; ++x) o.idx[a * n + x] = static_cast<std::uint32_t>(idx(a, x));
    }
    if (w) {
      o.w.resize(n * n);
      for (Z a = 0; a < n; ++a)
        for (Z x = 0; x < n; ++x) o.w[a * n + x] = w(a, x);
    }
    return o;
  }
  IdxFn tx() const {
    const Z nn = n;
    return [nn](Z a, Z x) { return a * nn + x; };
  }
  IdxFn xt() const {
    const Z nn = n;
    return [nn](Z a, Z x) { return x * nn + a; };
  }
  std::vector<T> U() const { return std::vector<T>(naux * n, T{}); }
  static std::vector<T> neg(const std::vector<T>& a) {
    std::vector<T> r(a.size());
    for (std::size_t i = 0; i < a.size(); ++i) r[i] = -a[i];
    return r;
  }
  std::vector<double> transposed(const double* m) const {
    std::vector<double> r(n * n);
    for (Z a = 0; a < n; ++a)
      for (Z b = 0; b < n; ++b) r[a * n + b] = m[b * n + a];
    return r;
  }
  void pc(std::vector<T>& tg, double sign, const Op<T>& f, const Op<T>& g) const { pairContract(tg, sign, n, naux, f, g); }

  // JkOnlySigmaHessianVectorRi.cpp:337-410
  void jkPhi(std::vector<T>& Phi) const {
    const T* Bm = t.b;
    const T* K1 = t.kb1.data();
    const T* K2 = t.kb2.data();
    for (Z p = 0; p < n; ++p)
      for (Z q = 0; q < n; ++q) {
        T kh{}, hk{};
        for (Z k = 0; k < n; ++k) {
          kh += KK(q, k) * h[k * n + p];
          hk += h[q * n + k] * KK(k, p);
        }
        Phi[p * n + q] += (occ[p] - occ[q]) * (kh - hk);
      }
    std::vector<T> Uh = U();
    accumU(Uh, 1.0, n, naux, t.bdiag.data(), hc);
    const std::vector<T> nUh = neg(Uh);
    diagContract(Phi, n, naux, K1, Uh.data(), nUh.data());
    pc(Phi, -1.0, op(Bm, tx(), [&](Z i, Z p) { return XC(p, i); }), op(K1, xt()));
    pc(Phi, 1.0, op(Bm, tx()), op(K1, xt(), [&](Z i, Z q) { return XC(q, i); }));
    diagContract(Phi, n, naux, K2, nUh.data(), Uh.data());
    pc(Phi, 1.0, op(K2, tx(), [&](Z i, Z p) { return XC(p, i); }), op(Bm, xt()));
    pc(Phi, -1.0, op(K2, tx()), op(Bm, xt(), [&](Z i, Z q) { return XC(q, i); }));
    {
      std::vector<T> V = U();
      accumU(V, 1.0, n, naux, t.kb1diag.data(), hc);
      accumU(V, -1.0, n, naux, t.kb2diag.data(), hc);
      const std::vector<T> nV = neg(V);
      diagContract(Phi, n, naux, Bm, V.data(), nV.data());
    }
    pc(Phi, 1.0, op(Bm, tx(), [&](Z i, Z p) { return XC(p, i); }), op(K2, xt()));
    pc(Phi, -1.0, op(Bm, tx()), op(K2, xt(), [&](Z i, Z q) { return XC(q, i); }));
    pc(Phi, -1.0, op(K1, tx(), [&](Z i, Z p) { return XC(p, i); }), op(Bm, xt()));
    pc(Phi, 1.0, op(K1, tx()), op(Bm, xt(), [&](Z i, Z q) { return XC(q, i); }));
  }

  // JkOnlySigmaHessianVectorRi.cpp:413-500
  void jkPsi(std::vector<T>& Psi) const {
    const T* Bm = t.b;
    const T* K1 = t.kb1.data();
    const T* K2 = t.kb2.data();
    for (Z c = 0; c < n; ++c)
      for (Z d = 0; d < n; ++d) {
        T acc{};
        for (Z a = 0; a < n; ++a) acc += KK(a, c) * (occ[a] - occ[c]) * h[d * n + a];
        for (Z b = 0; b < n; ++b) acc -= KK(d, b) * (occ[d] - occ[b]) * h[b * n + c];
        Psi[c * n + d] += acc;
      }
    std::vector<T> Uh = U();
    accumU(Uh, 1.0, n, naux, t.bdiag.data(), hc);
    const std::vector<T> nUh = neg(Uh);
    {
      Op<T> f = op(nullptr, nullptr, [&](Z a, Z c) { return KK(a, c); });
      f.row = Uh.data();
      pc(Psi, 1.0, f, op(Bm, xt()));
    }
    diagContract(Psi, n, naux, K2, nUh.data(), static_cast<const T*>(nullptr));
    {
      Op<T> f = op(Bm, tx(), [&](Z i, Z a) { return XC(a, i); });
      f.post.assign(K, K + n * n);
      pc(Psi, -1.0, f, op(Bm, xt()));
    }
    pc(Psi, 1.0, op(K2, tx(), [&](Z i, Z c) { return XC(c, i); }), op(Bm, xt()));
    diagContract(Psi, n, naux, K1, static_cast<const T*>(nullptr), nUh.data());
    {
      Op<T> f = op(Bm, tx());
      f.row = Uh.data();
      pc(Psi, 1.0, f, op(nullptr, nullptr, [&](Z b, Z d) { return KK(d, b); }));
    }
    pc(Psi, 1.0, op(Bm, tx()), op(K1, xt(), [&](Z i, Z d) { return XC(d, i); }));
    {
      Op<T> g = op(Bm, xt(), [&](Z i, Z b) { return XC(b, i); });
      g.post = transposed(K);
      pc(Psi, -1.0, op(Bm, tx()), g);
    }
    {
      const std::vector<double> hcT = transposed(hc);
      std::vector<T> V = U();
      accumU(V, 1.0, n, naux, t.kb1diag.data(), hcT.data());
      accumU(V, -1.0, n, naux, t.kb2diag.data(), hcT.data());
      const std::vector<T> nV = neg(V);
      diagContract(Psi, n, naux, Bm, V.data(), nV.data());
    }
    pc(Psi, 1.0, op(K1, tx()), op(Bm, xt(), [&](Z i, Z d) { return XC(i, d); }));
    pc(Psi, -1.0, op(K1, tx(), [&](Z i, Z c) { return XC(i, c); }), op(Bm, xt()));
    pc(Psi, 1.0, op(Bm, tx(), [&](Z i, Z c) { return XC(i, c); }), op(K2, xt()));
    pc(Psi, -1.0, op(Bm, tx()), op(K2, xt(), [&](Z i, Z d) { return XC(i, d); }));
  }

  // HartreeExchangeSigmaHessianVectorRi.cpp:220-293
  void hxPhi(std::vector<T>& Phi) const {
    const T* Bm = t.b;
    const T* K1 = t.kb1.data();
    const T* K2 = t.kb2.data();
    for (Z a = 0; a < n; ++a)
      for (Z b = 0; b < n; ++b) {
        T acc{}, t3{}, t4{};
        for (Z s = 0; s < n; ++s) acc += KK(b, s) * fock[a * n + s];
        for (Z r = 0; r < n; ++r) acc += conjv(fock[b * n + r]) * KK(r, a);
        for (Z s = 0; s < n; ++s) t3 += KK(b, s) * h[s * n + a];
        for (Z r = 0; r < n; ++r) t4 += h[b * n + r] * KK(r, a);
        Phi[a * n + b] += acc - occ[b] * t3 - occ[a] * t4;
      }
    std::vector<T> Uh = U();
    accumU(Uh, 1.0, n, naux, t.bdiag.data(), hc);
    const std::vector<T> nUh = neg(Uh);
    pc(Phi, 1.0, op(Bm, tx()), op(K1, xt(), [&](Z u, Z b) { return plantedSigma("hx_h_as_x") ? XC(b, u) : HC(b, u); }));
    diagContract(Phi, n, naux, K1, static_cast<const T*>(nullptr), nUh.data());
    pc(Phi, 1.0, op(K2, tx(), [&](Z u, Z a) { return HC(a, u); }), op(Bm, xt()));
    diagContract(Phi, n, naux, K2, nUh.data(), static_cast<const T*>(nullptr));
    pc(Phi, 1.0, op(K1, tx()), op(Bm, xt(), [&](Z r, Z b) { return XC(r, b); }));
    pc(Phi, -1.0, op(K1, tx(), [&](Z r, Z a) { return XC(r, a); }), op(Bm, xt()));
    pc(Phi, 1.0, op(Bm, tx(), [&](Z s, Z a) { return XC(a, s); }), op(K2, xt()));
    pc(Phi, -1.0, op(Bm, tx()), op(K2, xt(), [&](Z s, Z b) { return HC(s, b); }));
    const std::vector<double> hcT = transposed(hc), xcT = transposed(xc);
    std::vector<T> Ua = U(), Ub = U();
    accumU(Ua, 1.0, n, naux, t.kb1diag.data(), hcT.data());
    accumU(Ua, -1.0, n, naux, t.kb2diag.data(), xc);
    accumU(Ub, -1.0, n, naux, t.kb1diag.data(), xcT.data());
    accumU(Ub, 1.0, n, naux, t.kb2diag.data(), xc);
    diagContract(Phi, n, naux, Bm, Ua.data(), Ub.data());
  }

  // HartreeExchangeSigmaHessianVectorRi.cpp:295-363
  void hxPsi(std::vector<T>& Psi) const {
    const T* Bm = t.b;
    const T* K1 = t.kb1.data();
    const T* K2 = t.kb2.data();
    for (Z c = 0; c < n; ++c)
      for (Z d = 0; d < n; ++d) {
        T acc{}, t3{}, t4{};
        for (Z a = 0; a < n; ++a) acc += KK(a, c) * fock[a * n + d];
        for (Z b = 0; b < n; ++b) acc += KK(d, b) * conjv(fock[b * n + c]);
        for (Z a = 0; a < n; ++a) t3 += KK(a, c) * h[d * n + a];
        for (Z b = 0; b < n; ++b) t4 += KK(d, b) * h[b * n + c];
        Psi[c * n + d] += acc - occ[c] * t3 - occ[d] * t4;
      }
    std::vector<T> Uh = U();
    accumU(Uh, 1.0, n, naux, t.bdiag.data(), hc);
    const std::vector<T> nUh = neg(Uh);
    pc(Psi, 1.0, op(K2, tx(), [&](Z u, Z c) { return HC(c, u); }), op(Bm, xt()));
    diagContract(Psi, n, naux, K2, nUh.data(), static_cast<const T*>(nullptr));
    pc(Psi, 1.0, op(Bm, tx()), op(K1, xt(), [&](Z u, Z d) { return HC(d, u); }));
    diagContract(Psi, n, naux, K1, static_cast<const T*>(nullptr), nUh.data());
    pc(Psi, 1.0, op(Bm, tx(), [&](Z b, Z c) { return XC(c, b); }), op(K2, xt()));
    pc(Psi, -1.0, op(Bm, tx()), op(K2, xt(), [&](Z b, Z d) { return HC(d, b); }));
    pc(Psi, 1.0, op(K1, tx()), op(Bm, xt(), [&](Z a, Z d) { return XC(a, d); }));
    pc(Psi, -1.0, op(K1, tx(), [&](Z a, Z c) { return XC(c, a); }), op(Bm, xt()));
    const std::vector<double> xcT = transposed(xc);
    std::vector<T> Ua = U(), Ub = U();
    accumU(Ua, 1.0, n, naux, t.kb1diag.data(), hc);
    accumU(Ua, -1.0, n, naux, t.kb2diag.data(), xc);
    accumU(Ub, -1.0, n, naux, t.kb1diag.data(), xcT.data());
    accumU(Ub, 1.0, n, naux, t.kb2diag.data(), xcT.data());
    diagContract(Psi, n, naux, Bm, Ua.data(), Ub.data());
  }

  // HartreeExchangeSigmaHessianVectorRi.cpp:365-400
  void l12Phi(std::vector<T>& Phi) const {
    const T* Bm = t.b;
    const T* K1 = t.kb1.data();
    const T* K2 = t.kb2.data();
    const Z nn = n;
    pc(Phi, -1.0, op(Bm, tx()),
       op(K2, [&](Z u, Z b) { return pr(u) * nn + pr(b); },
          [&](Z u, Z b) { return 2.0 * (L1(u, pr(b)) + (plantedSigma("hx_l2_flip") ? L2(b, u) : L2(u, b))); }));
    pc(Phi, 1.0, op(Bm, [&](Z r, Z a) { return pr(r) * nn + a; }),
       op(K1, [&](Z r, Z b) { return r * nn + pr(b); }, [&](Z r, Z b) { return 2.0 * (L1(r, b) + L2(r, pr(b))); }));
    pc(Phi, -1.0, op(K1, tx()),
       op(Bm, [&](Z r, Z b) { return pr(r) * nn + pr(b); }, [&](Z r, Z b) { return 2.0 * (L1(r, b) + L2(r, pr(b))); }));
    pc(Phi, 1.0,
       op(K2, [&](Z s, Z a) { return pr(a) * nn + s; }, [&](Z s, Z a) { return 2.0 * (L1(a, s) + L2(a, pr(s))); }),
       op(Bm, [&](Z s, Z b) { return b * nn + pr(s); }));
    pc(Phi, -1.0,
       op(Bm, [&](Z s, Z a) { return pr(a) * nn + pr(s); }, [&](Z s, Z a) { return 2.0 * (L1(a, s) + L2(a, pr(s))); }),
       op(K2, xt()));
    pc(Phi, 1.0,
       op(K1, [&](Z w, Z a) { return pr(a) * nn + w; }, [&](Z w, Z a) { return 2.0 * (L1(pr(a), w) + L2(pr(a), pr(w))); }),
       op(Bm, [&](Z w, Z b) { return b * nn + pr(w); }));
  }

  // HartreeExchangeSigmaHessianVectorRi.cpp:402-438
  void l12Psi(std::vector<T>& Psi) const {
    const T* Bm = t.b;
    const T* K1 = t.kb1.data();
    const T* K2 = t.kb2.data();
    const Z nn = n;
    pc(Psi, -1.0, op(Bm, [&](Z u, Z c) { return pr(u) * nn + c; }),
       op(K2, [&](Z u, Z d) { return u * nn + pr(d); }, [&](Z u, Z d) { return 2.0 * (L1(u, d) + L2(u, pr(d))); }));
    pc(Psi, 1.0,
       op(K2, [&](Z b, Z c) { return pr(c) * nn + b; }, [&](Z b, Z c) { return 2.0 * (L1(c, b) + L2(c, pr(b))); }),
       op(Bm, [&](Z b, Z d) { return d * nn + pr(b); }));
    pc(Psi, -1.0,
       op(Bm, [&](Z b, Z c) { return pr(c) * nn + pr(b); }, [&](Z b, Z c) { return 2.0 * (L1(c, b) + L2(c, pr(b))); }),
       op(K2, xt()));
    pc(Psi, 1.0, op(Bm, [&](Z a, Z c) { return pr(a) * nn + c; }),
       op(K1, [&](Z a, Z d) { return a * nn + pr(d); }, [&](Z a, Z d) { return 2.0 * (L1(a, d) + L2(a, pr(d))); }));
    pc(Psi, -1.0, op(K1, tx()),
       op(Bm, [&](Z a, Z d) { return pr(a) * nn + pr(d); }, [&](Z a, Z d) { return 2.0 * (L1(a, d) + L2(a, pr(d))); }));
    pc(Psi, 1.0,
       op(K1, [&](Z w, Z c) { return pr(c) * nn + pr(w); }, [&](Z w, Z c) { return 2.0 * (L1(c, w) + L2(c, pr(w))); }),
       op(Bm, xt()));
  }
};

template <typename T>
void phiPsi(bool hx, std::size_t n, std::size_t naux, const T* b, const T* h, const double* occ, const double* hc,
            const double* xc, const T* fock, const std::int64_t* pair, const double* l1, const double* l2,
            const std::vector<double>& K, std::vector<T>& Phi, std::vector<T>& Psi) {
  const Toolkit<T> t = buildToolkit<T>(n, naux, b, K.data());
  Sigma<T> s{n, naux, t, h, occ, hc, xc, K.data(), fock, pair, l1, l2};
  Phi.assign(n * n, T{});
  Psi.assign(n * n, T{});
  if (!hx) {
    s.jkPhi(Phi);
    s.jkPsi(Psi);
  } else {
    s.hxPhi(Phi);
    s.hxPsi(Psi);
    if (pair) {
      s.l12Phi(Phi);
      s.l12Psi(Psi);
    }
  }
}

template <typename T>
Op<T> readOp(std::size_t n, const T* s, const std::uint32_t* idx, const double* w, const T* row, const double* post) {
  Op<T> o;
  o.s = s;
  if (s) o.idx.assign(idx, idx + n * n);
  if (w) o.w.assign(w, w + n * n);
  o.row = row;
  if (post) o.post.assign(post, post + n * n);
  return o;
}

}  // namespace

extern "C" {

const char* oracle_sigma_last_error() { return g_err_sigma.c_str(); }

int oracle_ri_hx_fock_gemm(int cplx, std::size_t n, std::size_t naux, const double* b, const double* h,
                           const double* occ, const double* h2, const double* x2, const std::int64_t* pair,
                           const double* l1, const double* l2, double* f) {
  try {
    if (cplx)
      hxFock<C>(n, naux, as<C>(b), as<C>(h), occ, h2, x2, pair, l1, l2, as<C>(f));
    else
      hxFock<double>(n, naux, b, h, occ, h2, x2, pair, l1, l2, f);
    return 0;
  } catch (const std::exception& e) {
    g_err_sigma = e.what();
    return 1;
  }
}

int oracle_ri_sigma_toolkit(int cplx, std::size_t n, std::size_t naux, const double* b, const double* K, double* kb1,
                            double* kb2, double* bdiag, double* kb1diag, double* kb2diag) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    const Toolkit<T> t = buildToolkit<T>(n, naux, as<T>(b), K);
    std::copy(t.kb1.begin(), t.kb1.end(), as<T>(kb1));
    std::copy(t.kb2.begin(), t.kb2.end(), as<T>(kb2));
    std::copy(t.bdiag.begin(), t.bdiag.end(), as<T>(bdiag));
    std::copy(t.kb1diag.begin(), t.kb1diag.end(), as<T>(kb1diag));
    std::copy(t.kb2diag.begin(), t.kb2diag.end(), as<T>(kb2diag));
  };
  if (cplx)
    run(C{});
  else
    run(0.0);
  return 0;
}

int oracle_ri_pair_contract(int cplx, std::size_t n, std::size_t naux, double sign, const double* fs,
                            const std::uint32_t* fidx, const double* fw, const double* frow, const double* fpost,
                            const double* gs, const std::uint32_t* gidx, const double* gw, const double* grow,
                            const double* gpost, double* target) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    std::vector<T> tg(as<T>(target), as<T>(target) + n * n);
    pairContract(tg, sign, n, naux, readOp<T>(n, as<T>(fs), fidx, fw, as<T>(frow), fpost),
                 readOp<T>(n, as<T>(gs), gidx, gw, as<T>(grow), gpost));
    std::copy(tg.begin(), tg.end(), as<T>(target));
  };
  if (cplx)
    run(C{});
  else
    run(0.0);
  return 0;
}

int oracle_ri_diag_contract(int cplx, std::size_t n, std::size_t naux, const double* s, const double* ua,
                            const double* ub, double* target) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    std::vector<T> tg(as<T>(target), as<T>(target) + n * n);
    diagContract<T>(tg, n, naux, as<T>(s), as<T>(ua), as<T>(ub));
    std::copy(tg.begin(), tg.end(), as<T>(target));
  };
  if (cplx)
    run(C{});
  else
    run(0.0);
  return 0;
}

int oracle_ri_accum_u(int cplx, std::size_t n, std::size_t naux, double alpha, const double* d, const double* wm,
                      double* u) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    std::vector<T> uu(as<T>(u), as<T>(u) + naux * n);
    accumU<T>(uu, alpha, n, naux, as<T>(d), wm);
    std::copy(uu.begin(), uu.end(), as<T>(u));
  };
  if (cplx)
    run(C{});
  else
    run(0.0);
  return 0;
}

// jkOnlySigmaHessianVectorRi / ...Joint... (JkOnlySigmaHessianVectorRi.cpp:504-573) and
// hartreeExchangeSigmaHessianVectorRi / ...Joint... (HartreeExchangeSigmaHessianVectorRi.cpp:442-539).
int oracle_sigma_hessian_vector_ri(int cplx, int hx, std::size_t n, const double* h, const double* b, std::size_t naux,
                                   const double* occ, const double* hc, const double* xc, const double* fock,
                                   const std::int64_t* pair, const double* l1, const double* l2, std::size_t npairs,
                                   const std::size_t* pairs, const double* v, double* w) {
  try {
    if (!cplx) {
      std::vector<double> K(n * n, 0.0);
      for (std::size_t i = 0; i < npairs; ++i) {
        K[pairs[2 * i] * n + pairs[2 * i + 1]] = v[i];
        K[pairs[2 * i + 1] * n + pairs[2 * i]] = -v[i];
      }
      std::vector<double> Phi, Psi;
      phiPsi<double>(hx != 0, n, naux, b, h, occ, hc, xc, fock, pair, l1, l2, K, Phi, Psi);
      for (std::size_t i = 0; i < npairs; ++i) {
        const std::size_t p = pairs[2 * i], q = pairs[2 * i + 1];
        w[i] = 0.5 * ((Phi[p * n + q] - Phi[q * n + p]) + (Psi[p * n + q] - Psi[q * n + p]));
      }
      return 0;
    }
    std::vector<double> Kt(n * n, 0.0), Ky(n * n, 0.0);
    for (std::size_t i = 0; i < npairs; ++i) {
      const std::size_t p = pairs[2 * i], q = pairs[2 * i + 1];
      Kt[p * n + q] = v[i];
      Kt[q * n + p] = -v[i];
      Ky[p * n + q] = v[npairs + i];
      Ky[q * n + p] = v[npairs + i];
    }
    std::vector<C> phit, psit, phiy, psiy;
    phiPsi<C>(hx != 0, n, naux, as<C>(b), as<C>(h), occ, hc, xc, as<C>(fock), pair, l1, l2, Kt, phit, psit);
    phiPsi<C>(hx != 0, n, naux, as<C>(b), as<C>(h), occ, hc, xc, as<C>(fock), pair, l1, l2, Ky, phiy, psiy);
    const C im(0.0, 1.0);
    for (std::size_t i = 0; i < npairs; ++i) {
      const std::size_t p = pairs[2 * i], q = pairs[2 * i + 1], pq = p * n + q, qp = q * n + p;
      w[i] = (0.5 * ((phit[pq] - phit[qp]) + (psit[pq] - psit[qp]))).real() +
             (0.5 * im * ((phiy[pq] - phiy[qp]) + (psiy[pq] - psiy[qp]))).real();
      w[npairs + i] = (0.5 * im * ((psit[pq] + psit[qp]) + (phit[pq] + phit[qp]))).real() -
                      (0.5 * ((phiy[pq] + phiy[qp]) + (psiy[pq] + psiy[qp]))).real();
    }
    return 0;
  } catch (const std::exception& e) {
    g_err_sigma = e.what();
    return 1;
  }
}

}  // extern "C"
