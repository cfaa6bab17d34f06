// Replays the call sequence of bench.py through the validation build of librerdmft_gpu
// (-DRR_VALIDATE, linked against tests/validate/hip_mock.cpp) on the CPU: the large real workload
// (20 incremental builds between full builds), the uneven batch, the shared-tensor sigma leg, the
// X2C / RKB legs and then the per-row leg, in that order and at the benchmark's sizes, with every
// caller array allocated at the extent bench.py gives it. No kernel runs; every launch, copy and
// solver call is checked by the validation hooks. TEST INFRASTRUCTURE ONLY.
//
// Usage: replay [legs]   legs: any of "main,batch,sigma,rel,rows,noalloc" (default: all, in that order);
// "rows" replays the per-row leg as bench.py calls it; "plant" runs planted caller mistakes. Exit status: 0 when no violation and no
// failed call, 1 otherwise. The number of violations is printed last.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/rerdmft_gpu.h"

extern "C" long rrv_violations();
extern "C" long rrv_checks();

namespace {

rr_handle H = nullptr;
int failed_calls = 0;

double* dev(size_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double)) != hipSuccess) {
    std::fprintf(stderr, "replay: allocation of %zu doubles failed\n", count);
    std::exit(2);
  }
  return static_cast<double*>(p);
}
void release(double* p) { (void)hipFree(p); }

void call(const char* name, int rc) {
  if (rc != RR_OK) {
    ++failed_calls;
    std::printf("  %-44s rc=%d: %s\n", name, rc, rr_last_error(H));
  } else {
    std::printf("  %-44s ok (violations so far %ld)\n", name, rrv_violations());
  }
  std::fflush(stdout);
}

void control(const char* spec) { setenv("RRV_CONTROL", spec, 1); }

// Recovery of an occupied set of `occ` vectors through the spectral projector.
std::string projector(size_t occ) {
  return "projector_scale=1,projector_residual=0,projector_trace=" + std::to_string(occ);
}

constexpr unsigned ALL = RR_COPY_FP64 | RR_COPY_FP32 | RR_COPY_BF16;
constexpr uint64_t SEED = 20261203 + 1;

void leg_main() {
  std::printf("main: B[4000][1000][1000], n_occ = 200, 20 incremental builds + full builds\n");
  const size_t n = 1000, naux = 4000, no = 200;
  call("rr_tensor_generate slot 0", rr_tensor_generate(H, 0, naux, n, 1.0, 50.0, SEED, ALL));
  std::vector<double*> c(21), l(21);
  for (int k = 0; k <= 20; ++k) c[k] = dev(n * no), l[k] = dev(n * no);
  double *J = dev(n * n), *K = dev(n * n), *dJ = dev(n * n), *dK = dev(n * n);
  auto full = [&](int k, int prec) {
    call("rr_ri_jk_factor (fused J)", rr_ri_jk_factor(H, 0, prec, n, nullptr, l[k], nullptr, no, J, K));
  };
  control("align_scale=1,align_residual=0");  // the polar alignment of the incremental builds converges
  for (int rep = 0; rep < 2; ++rep) {  // warm-up and one timed sequence
    full(0, RR_FP64);
    for (int k = 1; k <= 20; ++k)
      call("rr_ri_jk_delta bf16", rr_ri_jk_delta(H, 0, RR_BF16, n, no, c[k], c[k - 1], dJ, dK));
    full(20, RR_FP64);
  }
  control("");
  for (int k = 0; k < 22; ++k) full(std::min(k, 20), RR_FP64);
  for (int k = 0; k <= 20; ++k) release(c[k]), release(l[k]);
  release(J), release(K), release(dJ), release(dK);
}

void leg_batch() {
  std::printf("batch: 256 molecules, n ~ U[20,200], n_aux = 4n, n_occ = n/5\n");
  std::mt19937_64 g(20261203 + 5);
  std::uniform_int_distribution<int> u(20, 200);
  const size_t nm = 256;
  std::vector<size_t> ns(nm), naux(nm), r(nm);
  std::vector<const double*> tens(nm), ls(nm);
  std::vector<double*> js(nm), ks(nm), own;
  for (size_t m = 0; m < nm; ++m) {
    ns[m] = (size_t)u(g), naux[m] = 4 * ns[m], r[m] = std::max<size_t>(1, ns[m] / 5);
    double* t = dev(naux[m] * ns[m] * ns[m]);
    own.push_back(t);
    tens[m] = t;
  }
  rr_batch_t batch = nullptr;
  call("rr_batch_create", rr_batch_create(H, nm, naux.data(), ns.data(), tens.data(), RR_COPY_FP64, &batch));
  for (double* t : own) release(t);
  for (size_t m = 0; m < nm; ++m) {
    double* lm = dev(ns[m] * r[m]);
    ls[m] = lm;
    js[m] = dev(ns[m] * ns[m]), ks[m] = dev(ns[m] * ns[m]);
  }
  for (int rep = 0; rep < 7; ++rep) {  // eager + capture, then graph replays
    int replayed = 0;
    call("rr_batch_jk_factor", rr_batch_jk_factor(H, batch, RR_FP64, r.data(), ls.data(), js.data(), ks.data(), &replayed));
  }
  call("rr_batch_destroy", rr_batch_destroy(H, batch));
  for (size_t m = 0; m < nm; ++m) release(const_cast<double*>(ls[m])), release(js[m]), release(ks[m]);
}

void leg_sigma() {
  std::printf("sigma: 64 trial densities, n = 800, n_occ = 100, n_aux = 3200\n");
  const size_t n = 800, no = 100, naux = 3200, nd = 64;
  call("rr_tensor_generate slot 14", rr_tensor_generate(H, 14, naux, n, 1.0, 40.0, SEED + 7, ALL));
  double *cocc = dev(n * no), *rf = dev(nd * n * no), *js = dev(nd * n * n), *ks = dev(nd * n * n);
  for (int prec : {RR_FP64, RR_FP32, RR_BF16})
    for (int rep = 0; rep < 3; ++rep) {
      int replayed = 0;
      call("rr_ri_jk_multi", rr_ri_jk_multi(H, 14, prec, n, no, cocc, nd, rf, js, ks, &replayed));
    }
  call("rr_tensor_release slot 14", rr_tensor_release(H, 14));
  release(cocc), release(rf), release(js), release(ks);
}

void leg_relativistic() {
  std::printf("relativistic: X2C n = 400 (n_aux 2400, 160 occupied) and MO transform; RKB nL 300 / nS 700 / n_aux 4200\n");
  {
    const size_t nl = 400, naux = 2400, occ = 160, nt = 800;
    call("rr_tensor_generate slot 13", rr_tensor_generate(H, 13, naux, nl, 1.0, 40.0, 20261203 + 2, ALL));
    double *eye = dev(2 * nt * nt), *p = dev(2 * nt * nt), *h = dev(2 * nt * nt), *f = dev(2 * nt * nt);
    control(projector(occ).c_str());
    for (int prec : {RR_FP64, RR_FP32, RR_BF16})
      for (int rep = 0; rep < 4; ++rep)
        call("rr_x2c_fock_ri", rr_x2c_fock_ri(H, 13, prec, nl, h, eye, nt, eye, p, f));
    double* u = dev(2 * nt * nt);
    for (int prec : {RR_FP64, RR_FP32}) {
      double* mo = dev((prec == RR_FP64 ? 2 : 1) * naux * nt * nt);
      for (int rep = 0; rep < 2; ++rep)
        call("rr_build_ri_mo_eri", rr_build_ri_mo_eri(H, 13, 13, prec, prec, nl, 0, u, nt, mo));
      release(mo);
    }
    call("rr_tensor_release slot 13", rr_tensor_release(H, 13));
    release(eye), release(p), release(h), release(f), release(u);
  }
  {
    const size_t nl = 300, ns = 700, naux = 4200, nll = 1800, no = 200, n = 2 * nl + 2 * ns;
    double* bll = dev(naux * nl * nl);
    control("upload_nonzero_rows=1800");
    call("rr_tensor_upload slot 10 (LL, zero past 1800)", rr_tensor_upload(H, 10, bll, naux, nl, ALL));
    (void)nll;
    release(bll);
    call("rr_tensor_generate slot 11", rr_tensor_generate(H, 11, naux, ns, 1.0, 70.0, 20261203 + 13, ALL));
    double *dens = dev(2 * n * n), *cq = dev(2 * n * no), *g = dev(2 * n * n);
    std::vector<double> occ(no, 1.0);  // host array, as bench.py passes numpy occupations
    control("");
    for (int prec : {RR_FP64, RR_FP32, RR_BF16})
      for (int rep = 0; rep < 4; ++rep)
        call("rr_ri_fock_two_electron_direct",
             rr_ri_fock_two_electron_direct(H, 10, 11, prec, nl, ns, dens, cq, no, occ.data(), g));
    call("rr_tensor_release slot 10", rr_tensor_release(H, 10));
    call("rr_tensor_release slot 11", rr_tensor_release(H, 11));
    release(dens), release(cq), release(g);
  }
}

void leg_rows() {
  std::printf("rows: the other entry points at mid-size shapes (as bench.py's per-row leg calls them)\n");
  {  // a4
    const size_t nl = 150, ns = 350, naux = 1200, nocc = 120, nu = 2 * nl + 2 * ns, nr = 4 * nl;
    call("rr_tensor_generate slot 5", rr_tensor_generate(H, 5, naux, nl, 0.4, 20.0, SEED + 31, RR_COPY_FP64));
    call("rr_tensor_generate slot 6", rr_tensor_generate(H, 6, naux, ns, 0.2, 40.0, SEED + 32, RR_COPY_FP64));
    double *h = dev(2 * nr * nr), *v = dev(2 * nu * nr), *sr = dev(2 * nr * nr), *x = dev(2 * nr * nr),
           *pd = dev(2 * nr * nr), *fo = dev(2 * nr * nr);
    control(projector(nocc).c_str());
    for (int rep = 0; rep < 4; ++rep)
      call("rr_rkb_fock_ri (a4)", rr_rkb_fock_ri(H, 5, 6, RR_FP64, nl, ns, h, v, nr, sr, x, nr, pd, fo));
    control("");
    call("rr_tensor_release slot 5", rr_tensor_release(H, 5));
    call("rr_tensor_release slot 6", rr_tensor_release(H, 6));
    release(h), release(v), release(sr), release(x), release(pd), release(fo);
  }
  {  // a5 / a15 / a8 / a10 on one set of n x n vectors
    const size_t n = 500, nchol = 1600, nmo = 250;
    call("rr_tensor_generate slot 7", rr_tensor_generate(H, 7, nchol, n, 1.0, 50.0, SEED + 41, RR_COPY_FP64));
    double *h = dev(n * n), *pdn = dev(n * n), *fo = dev(n * n);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_nonrel_fock_cholesky (a5)", rr_nonrel_fock_cholesky(H, 7, RR_FP64, n, h, pdn, fo));
    double *hx = dev(2 * 4 * n * n), *px = dev(2 * 4 * n * n), *fx = dev(2 * 4 * n * n);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_x2c_fock_cholesky (a5)", rr_x2c_fock_cholesky(H, 7, RR_FP64, n, hx, px, fx));
    double *cm = dev(n * nmo), *mo = dev(nchol * nmo * nmo);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_cholesky_to_mo_spatial (a15)", rr_cholesky_to_mo_spatial(H, 7, RR_FP64, n, cm, nmo, mo));
    for (int rep = 0; rep < 4; ++rep)
      call("rr_build_ri_nonrel_spin_mo_eri (a8)", rr_build_ri_nonrel_spin_mo_eri(H, 7, RR_FP64, n, cm, nmo, mo));
    double *occ = dev(2 * n), *h2 = dev(4 * n * n), *x2 = dev(4 * n * n), *fm = dev(4 * n * n);
    // c_current as bench.py's per-row leg allocates it
    const char* cshape = std::getenv("REPLAY_C_CURRENT_COLS");
    const size_t ccols = cshape ? (size_t)std::atol(cshape) : 2 * n;
    double* cfull = dev(n * ccols);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_ri_mo_fock_nonrel (a10)",
           rr_ri_mo_fock_nonrel(H, 7, n, h, cfull, occ, h2, x2, nullptr, nullptr, nullptr, fm));
    release(h), release(pdn), release(fo), release(hx), release(px), release(fx), release(cm), release(mo);
    release(occ), release(h2), release(x2), release(fm), release(cfull);
    call("rr_tensor_release slot 7", rr_tensor_release(H, 7));
  }
  {  // a9 / a13 on a device MO tensor
    const size_t n = 200, naux = 800;
    double *b = dev(naux * n * n), *hm = dev(n * n), *occ = dev(n), *h2 = dev(n * n), *x2 = dev(n * n), *fo = dev(n * n);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_ri_hx_fock_gemm (a9)",
           rr_ri_hx_fock_gemm(H, 0, n, naux, b, hm, occ, h2, x2, nullptr, nullptr, nullptr, fo));
    std::vector<size_t> pairs;
    for (size_t p = 0; p < n; ++p)
      for (size_t q = 0; q < p; ++q) pairs.push_back(p), pairs.push_back(q);
    const size_t np = pairs.size() / 2;
    double* vv = dev(np);
    std::vector<double> w(np);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_jk_only_sigma_hessian_vector_ri (a13)",
           rr_jk_only_sigma_hessian_vector_ri(H, n, hm, b, naux, occ, h2, x2, np, pairs.data(), vv, w.data()));
    release(b), release(hm), release(occ), release(h2), release(x2), release(fo), release(vv);
  }
  {  // a7
    const size_t nt = 400, naux = 600;
    double *bz = dev(2 * naux * nt * nt), *uz = dev(2 * nt * nt), *bo = dev(2 * naux * nt * nt);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_ri_mo_eri_rotated (a7)", rr_ri_mo_eri_rotated(H, RR_FP64, nt, naux, bz, uz, bo));
    release(bz), release(uz), release(bo);
  }
  {  // a14
    const size_t n = 200, nraw = 2000, kept = 1700;
    double *metric = dev(nraw * nraw), *xk = dev(nraw * nraw);
    size_t nk = 0;
    control(("eig_top=" + std::to_string(kept)).c_str());
    for (int rep = 0; rep < 4; ++rep)
      call("rr_canonical_orthogonalize (a14)", rr_canonical_orthogonalize(H, nraw, metric, 1e-10, xk, &nk));
    control("");
    double *xkk = dev(nraw * kept), *raw = dev(nraw * n * n), *bo = dev(kept * n * n);
    for (int rep = 0; rep < 4; ++rep)
      call("rr_ri_three_center_tensor (a14)", rr_ri_three_center_tensor(H, n, nraw, raw, xkk, kept, bo));
    release(metric), release(xk), release(xkk), release(raw), release(bo);
  }
}

// Planted caller mistakes the hooks must report (for the test of the validator itself): an output
// array freed before the call, and an input array shorter than the entry point's contract.
void leg_plant() {
  std::printf("plant: freed output / short input\n");
  const size_t n = 64, naux = 128, no = 8;
  call("rr_tensor_generate slot 1", rr_tensor_generate(H, 1, naux, n, 1.0, 8.0, SEED, RR_COPY_FP64));
  double *l = dev(n * no), *j = dev(n * n), *k = dev(n * n);
  release(k);
  call("rr_ri_jk_factor (freed K output)", rr_ri_jk_factor(H, 1, RR_FP64, n, nullptr, l, nullptr, no, j, k));
  double* lshort = dev(n * no - 1);
  k = dev(n * n);
  call("rr_ri_jk_factor (short factor)", rr_ri_jk_factor(H, 1, RR_FP64, n, nullptr, lshort, nullptr, no, j, k));
  release(l), release(lshort), release(j), release(k);
  call("rr_tensor_release slot 1", rr_tensor_release(H, 1));
}

// The allocation-free build: two calls in a row, the second smaller, on the static scratch region.
void leg_noalloc() {
  std::printf("noalloc: rr_ri_jk_factor_noalloc, n_aux 800 / n 200 / r 40, then 300 / 24 / 5\n");
  const size_t shapes[2][3] = {{800, 200, 40}, {300, 24, 5}};
  for (const auto& sh : shapes) {
    const size_t naux = sh[0], n = sh[1], r = sh[2];
    call("rr_tensor_generate slot 8", rr_tensor_generate(H, 8, naux, n, 1.0, 20.0, SEED, ALL));
    double *l = dev(n * r), *j = dev(n * n), *k = dev(n * n);
    for (int prec : {RR_FP64, RR_FP32, RR_BF16})
      call("rr_ri_jk_factor_noalloc", rr_ri_jk_factor_noalloc(H, 8, prec, n, l, r, j, k));
    release(l), release(j), release(k);
    call("rr_tensor_release slot 8", rr_tensor_release(H, 8));
  }
}

}  // namespace

int main(int argc, char** argv) {
  const std::string legs = argc > 1 ? argv[1] : "main,batch,sigma,rel,rows,noalloc";
  setvbuf(stdout, nullptr, _IOLBF, 0);
  if (rr_create(0, &H) != RR_OK) {
    std::fprintf(stderr, "rr_create failed\n");
    return 2;
  }
  auto has = [&](const char* s) { return legs.find(s) != std::string::npos; };
  if (has("main")) leg_main();
  if (has("batch")) leg_batch();
  if (has("sigma")) leg_sigma();
  if (has("rel")) leg_relativistic();
  if (has("rows")) leg_rows();
  if (has("noalloc")) leg_noalloc();
  if (has("plant")) leg_plant();
  rr_destroy(H);
  std::printf("replay: %ld checks, %d failed calls, %ld violations\n", rrv_checks(), failed_calls, rrv_violations());
  return (failed_calls == 0 && rrv_violations() == 0) ? 0 : 1;
}
