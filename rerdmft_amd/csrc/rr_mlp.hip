// librerdmft_gpu, SCF starting guess model: a fully connected FP32 network (input -> L hidden layers
// of width H with tanh -> linear output) with its forward pass, back-propagation of the
// mean-squared error, and Adam updates, all on the device. The matrix products run on the NT
// MFMA engine of rr_gemm.hpp (FP32 operands and accumulation); bias, activation, loss and
// optimizer steps are elementwise kernels. No reference counterpart: it predicts the converged
// density of a system from descriptors, replacing the core-Hamiltonian guess where it helps.
//
// Parameters of layer l (fan_in -> fan_out): W_l [fan_out][fan_in] row-major, b_l [fan_out].
// Activations are kept per layer for the backward pass: A_0 = X [batch][n_in], A_l [batch][width].
#include "rr_core.hpp"

struct rr_mlp {
  std::vector<int> dims;        // n_in, H, ..., H, n_out
  std::vector<size_t> w_off, b_off;
  size_t n_params = 0;
  float* params = nullptr;      // all W and b
  float* grads = nullptr;
  float* m1 = nullptr;          // Adam moments
  float* m2 = nullptr;
  long long step = 0;
  size_t batch_cap = 0;
  std::vector<float*> act;      // A_0..A_L (padded rows)
  std::vector<float*> dact;     // dZ per layer
  float* loss = nullptr;        // device scalar partials
};

namespace {

using rr::Context;
using rr::Error;

#define RR_GRID_LOOP(i, total) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (total); i += (size_t)gridDim.x * blockDim.x)

inline size_t ldp(int d) { return rr::roundup(static_cast<size_t>(std::max(d, 1)), 8); }

// splitmix64 -> uniform in (-a, a)
__global__ void k_mlp_init(float* p, size_t count, float a, unsigned long long seed) {
  RR_GRID_LOOP(i, count) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    p[i] = static_cast<float>((2.0 * u - 1.0) * a);
  }
}
__global__ void k_mlp_zero(float* p, size_t count) {
  RR_GRID_LOOP(i, count) p[i] = 0.f;
}
// dst [rows][ldd] (zero-padded columns) = src [rows][cols] (double, host layout)
__global__ void k_mlp_load_rows(const double* src, size_t rows, size_t cols, size_t ldd, float* dst) {
  RR_GRID_LOOP(i, rows * ldd) {
    const size_t r = i / ldd, c = i % ldd;
    dst[i] = c < cols ? static_cast<float>(src[r * cols + c]) : 0.f;
  }
}
// Padded weight copy: wp[o][ldi] = W[o][i]; transposed copy wt[i][ldo] = W[o][i]
__global__ void k_mlp_pack_w(const float* w, int fo, int fi, size_t ldi, size_t ldo, float* wp, float* wt) {
  const size_t total = static_cast<size_t>(fo) * ldi;
  RR_GRID_LOOP(e, total) {
    const size_t o = e / ldi, i = e % ldi;
    wp[e] = i < static_cast<size_t>(fi) ? w[o * fi + i] : 0.f;
  }
  const size_t tt = static_cast<size_t>(fi) * ldo;
  RR_GRID_LOOP(e, tt) {
    const size_t i = e / ldo, o = e % ldo;
    wt[e] = o < static_cast<size_t>(fo) ? w[o * fi + i] : 0.f;
  }
}
// Z [batch][ldo] += b (columns < fo), then A = tanh(Z) in place (hidden layers) or identity (output)
__global__ void k_mlp_bias_act(float* z, size_t batch, int fo, size_t ldo, const float* b, int act) {
  RR_GRID_LOOP(i, batch * ldo) {
    const size_t o = i % ldo;
    if (o < static_cast<size_t>(fo)) {
      const float v = z[i] + b[o];
      z[i] = act ? tanhf(v) : v;
    } else {
      z[i] = 0.f;
    }
  }
}
// dY = (Y - T) / (batch * n_out) and the loss partial sum of squares
__global__ __launch_bounds__(256) void k_mlp_loss(const float* y, const double* t, size_t batch, int no, size_t ld,
                                                  float* dy, float* loss) {
  __shared__ float sh[256];
  float acc = 0.f;
  const float scale = 1.f / static_cast<float>(batch * no);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < batch * ld; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / ld, c = i % ld;
    float d = 0.f;
    if (c < static_cast<size_t>(no)) {
      d = y[i] - static_cast<float>(t[r * no + c]);
      acc += d * d;
    }
    dy[i] = d * scale;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, 0.5f * sh[0] * scale);
}
// dZ = dA * (1 - A^2)  (tanh'), in place on dA; A is the layer output
__global__ void k_mlp_act_back(float* da, const float* a, size_t count) {
  RR_GRID_LOOP(i, count) da[i] *= (1.f - a[i] * a[i]);
}
// db[o] = sum_b dZ[b][o]
__global__ void k_mlp_bias_grad(const float* dz, size_t batch, int fo, size_t ldo, float* db) {
  RR_GRID_LOOP(o, static_cast<size_t>(fo)) {
    float s = 0.f;
    for (size_t b = 0; b < batch; ++b) s += dz[b * ldo + o];
    db[o] = s;
  }
}
// t[c][ldr] (zero-padded columns) = m[r][c] for r < rows, c < cols (m row length ldm)
__global__ void k_mlp_transpose(const float* m, size_t rows, size_t cols, size_t ldm, size_t ldr, float* t) {
  RR_GRID_LOOP(i, cols * ldr) {
    const size_t c = i / ldr, r = i % ldr;
    t[i] = r < rows ? m[r * ldm + c] : 0.f;
  }
}
// dW[o][i] (row length fi, unpadded) from a padded product gw[o][ldi]
__global__ void k_mlp_unpad(const float* gw, int fo, int fi, size_t ldi, float* dw) {
  RR_GRID_LOOP(e, static_cast<size_t>(fo) * fi) dw[e] = gw[(e / fi) * ldi + e % fi];
}
__global__ void k_mlp_adam(float* p, const float* g, float* m, float* v, size_t count, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
  RR_GRID_LOOP(i, count) {
    const float gi = g[i];
    m[i] = b1 * m[i] + (1.f - b1) * gi;
    v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
    p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
  }
}
__global__ void k_mlp_store_rows(const float* src, size_t rows, size_t cols, size_t lds, double* dst) {
  RR_GRID_LOOP(i, rows * cols) dst[i] = static_cast<double>(src[(i / cols) * lds + i % cols]);
}

// C[m][n] = sum_k A[m][k] B[n][k]  (FP32 NT, padded rows)
void gemm32(Context& c, const float* A, size_t lda, const float* B, size_t ldb, float* C, size_t ldc, int M, int N, int K) {
  rr::GemmArgs g;
  g.A = A, g.lda = (long long)lda;
  g.B = B, g.ldb = (long long)ldb;
  g.C = C, g.ldcm = (long long)ldc, g.ldcn = 1;
  g.M = M, g.N = N, g.K = K;
  rr::GemmSpec sp;
  sp.prec = rr::P32, sp.out = rr::P32, sp.cfg = 1;
  rr::run_gemm(c, sp, g);
}

enum MlpWs { WS_ML_WP = rr::WS_SIGMA_BASE + 4000, WS_ML_WT, WS_ML_T1, WS_ML_T2, WS_ML_GW, WS_ML_X, WS_ML_T, WS_ML_Y };

void ensure_batch(rr_mlp* m, size_t batch) {
  if (batch <= m->batch_cap) return;
  for (float* p : m->act) (void)hipFree(p);
  for (float* p : m->dact) (void)hipFree(p);
  m->act.assign(m->dims.size(), nullptr);
  m->dact.assign(m->dims.size(), nullptr);
  for (size_t l = 0; l < m->dims.size(); ++l) {
    RR_HIP(hipMalloc(&m->act[l], batch * ldp(m->dims[l]) * sizeof(float)));
    RR_HIP(hipMalloc(&m->dact[l], batch * ldp(m->dims[l]) * sizeof(float)));
    // the padding columns of dZ are read (times zero weights) by the next product: keep them zero
    RR_HIP(hipMemset(m->dact[l], 0, batch * ldp(m->dims[l]) * sizeof(float)));
  }
  m->batch_cap = batch;
}

// Forward pass; the output stays in act.back() (padded rows).
void forward(Context& c, rr_mlp* m, size_t batch, const double* dx) {
  const int L = static_cast<int>(m->dims.size()) - 1;
  hipLaunchKernelGGL(k_mlp_load_rows, dim3(rr::grid_for(batch * ldp(m->dims[0]))), dim3(256), 0, c.stream, dx, batch,
                     (size_t)m->dims[0], ldp(m->dims[0]), m->act[0]);
  for (int l = 0; l < L; ++l) {
    const int fi = m->dims[l], fo = m->dims[l + 1];
    const size_t li = ldp(fi), lo = ldp(fo);
    float* wp = static_cast<float*>(rr::wsbuf(c, WS_ML_WP + 16 * l, fo * li * sizeof(float)));
    float* wt = static_cast<float*>(rr::wsbuf(c, WS_ML_WT + 16 * l, fi * lo * sizeof(float)));
    hipLaunchKernelGGL(k_mlp_pack_w, dim3(rr::grid_for(std::max(fo * li, fi * lo))), dim3(256), 0, c.stream,
                       m->params + m->w_off[l], fo, fi, li, lo, wp, wt);
    gemm32(c, m->act[l], li, wp, li, m->act[l + 1], lo, (int)batch, fo, (int)li);  // Z = A W^T
    hipLaunchKernelGGL(k_mlp_bias_act, dim3(rr::grid_for(batch * lo)), dim3(256), 0, c.stream, m->act[l + 1], batch, fo,
                       lo, m->params + m->b_off[l], l + 1 < L ? 1 : 0);
    RR_HIP(hipGetLastError());
  }
}

// Loss and gradients (into m->grads) of 1/(2 B n_out) sum (y - t)^2.
void backward(Context& c, rr_mlp* m, size_t batch, const double* dt) {
  const int L = static_cast<int>(m->dims.size()) - 1;
  const int no = m->dims[L];
  RR_HIP(hipMemsetAsync(m->loss, 0, sizeof(float), c.stream));
  hipLaunchKernelGGL(k_mlp_loss, dim3((unsigned)std::min<size_t>(1024, (batch * ldp(no) + 255) / 256)), dim3(256), 0,
                     c.stream, m->act[L], dt, batch, no, ldp(no), m->dact[L], m->loss);
  for (int l = L - 1; l >= 0; --l) {
    const int fi = m->dims[l], fo = m->dims[l + 1];
    const size_t li = ldp(fi), lo = ldp(fo), lb = rr::roundup(batch, 8);
    float* dz = m->dact[l + 1];
    if (l + 1 < L)  // hidden layer: through tanh
      hipLaunchKernelGGL(k_mlp_act_back, dim3(rr::grid_for(batch * lo)), dim3(256), 0, c.stream, dz, m->act[l + 1],
                         batch * lo);
    hipLaunchKernelGGL(k_mlp_bias_grad, dim3(rr::grid_for(fo)), dim3(256), 0, c.stream, dz, batch, fo, lo,
                       m->grads + m->b_off[l]);
    // dW[o][i] = sum_b dZ[b][o] A[b][i]:  NT with dZ^T [o][b] and A^T [i][b]
    float* dzt = static_cast<float*>(rr::wsbuf(c, WS_ML_T1, lo * lb * sizeof(float)));
    float* at = static_cast<float*>(rr::wsbuf(c, WS_ML_T2, li * lb * sizeof(float)));
    hipLaunchKernelGGL(k_mlp_transpose, dim3(rr::grid_for(lo * lb)), dim3(256), 0, c.stream, dz, batch, lo, lo, lb, dzt);
    hipLaunchKernelGGL(k_mlp_transpose, dim3(rr::grid_for(li * lb)), dim3(256), 0, c.stream, m->act[l], batch, li, li, lb,
                       at);
    float* gw = static_cast<float*>(rr::wsbuf(c, WS_ML_GW, fo * li * sizeof(float)));
    gemm32(c, dzt, lb, at, lb, gw, li, fo, (int)li, (int)lb);
    hipLaunchKernelGGL(k_mlp_unpad, dim3(rr::grid_for((size_t)fo * fi)), dim3(256), 0, c.stream, gw, fo, fi, li,
                       m->grads + m->w_off[l]);
    if (l > 0) {  // dA_l [b][i] = sum_o dZ[b][o] W[o][i]:  NT with W^T [i][o]
      float* wt = static_cast<float*>(rr::wsbuf(c, WS_ML_WT + 16 * l, fi * lo * sizeof(float)));
      gemm32(c, dz, lo, wt, lo, m->dact[l], li, (int)batch, fi, (int)lo);
    }
    RR_HIP(hipGetLastError());
  }
}

}  // namespace

using rr::guarded;

extern "C" {

int rr_mlp_create(rr_handle h, int n_in, int n_hidden, int n_layers, int n_out, unsigned long long seed, rr_mlp_t* out) {
  if (!out) return RR_ERR_ARG;
  *out = nullptr;
  rr_mlp* m = new rr_mlp();
  const int rc = guarded(h, [&](Context& c) {
    RR_REQUIRE(n_in > 0 && n_hidden > 0 && n_layers > 0 && n_out > 0, RR_ERR_ARG, "mlp: sizes must be positive");
    m->dims.push_back(n_in);
    for (int l = 0; l < n_layers; ++l) m->dims.push_back(n_hidden);
    m->dims.push_back(n_out);
    for (size_t l = 0; l + 1 < m->dims.size(); ++l) {
      m->w_off.push_back(m->n_params);
      m->n_params += static_cast<size_t>(m->dims[l]) * m->dims[l + 1];
      m->b_off.push_back(m->n_params);
      m->n_params += m->dims[l + 1];
    }
    for (float** p : {&m->params, &m->grads, &m->m1, &m->m2}) RR_HIP(hipMalloc(p, m->n_params * sizeof(float)));
    RR_HIP(hipMalloc(&m->loss, sizeof(float)));
    for (size_t l = 0; l + 1 < m->dims.size(); ++l) {  // Glorot-uniform weights, zero biases
      const float a = std::sqrt(6.f / static_cast<float>(m->dims[l] + m->dims[l + 1]));
      const size_t nw = static_cast<size_t>(m->dims[l]) * m->dims[l + 1];
      hipLaunchKernelGGL(k_mlp_init, dim3(rr::grid_for(nw)), dim3(256), 0, c.stream, m->params + m->w_off[l], nw, a,
                         seed + 7919ull * l);
      hipLaunchKernelGGL(k_mlp_zero, dim3(rr::grid_for(m->dims[l + 1])), dim3(256), 0, c.stream, m->params + m->b_off[l],
                         (size_t)m->dims[l + 1]);
    }
    for (float* p : {m->grads, m->m1, m->m2})
      hipLaunchKernelGGL(k_mlp_zero, dim3(rr::grid_for(m->n_params)), dim3(256), 0, c.stream, p, m->n_params);
    RR_HIP(hipGetLastError());
  });
  if (rc != RR_OK) {
    for (float* p : {m->params, m->grads, m->m1, m->m2, m->loss}) if (p) (void)hipFree(p);
    delete m;
    return rc;
  }
  *out = m;
  return RR_OK;
}

int rr_mlp_destroy(rr_handle h, rr_mlp_t m) {
  if (!h || !m) return RR_ERR_ARG;
  return guarded(h, [&](Context& c) {
    RR_HIP(hipStreamSynchronize(c.stream));
    for (float* p : {m->params, m->grads, m->m1, m->m2, m->loss}) if (p) (void)hipFree(p);
    for (float* p : m->act) (void)hipFree(p);
    for (float* p : m->dact) (void)hipFree(p);
    delete m;
  });
}

int rr_mlp_num_params(rr_handle h, rr_mlp_t m, size_t* n) {
  if (!h || !m || !n) return RR_ERR_ARG;
  *n = m->n_params;
  return RR_OK;
}

int rr_mlp_get_params(rr_handle h, rr_mlp_t m, float* out) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(m && out, RR_ERR_ARG, "mlp: null argument");
    RR_HIP(hipMemcpyAsync(out, m->params, m->n_params * sizeof(float), hipMemcpyDefault, c.stream));
  });
}

int rr_mlp_set_params(rr_handle h, rr_mlp_t m, const float* in) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(m && in, RR_ERR_ARG, "mlp: null argument");
    RR_HIP(hipMemcpyAsync(m->params, in, m->n_params * sizeof(float), hipMemcpyDefault, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
  });
}

int rr_mlp_forward(rr_handle h, rr_mlp_t m, size_t batch, const double* x, double* y) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(m && x && y, RR_ERR_ARG, "mlp: null argument");
    if (batch == 0) return;
    ensure_batch(m, batch);
    const double* dx = rr::in_buf(c, WS_ML_X, x, batch * m->dims.front());
    forward(c, m, batch, dx);
    const int no = m->dims.back();
    rr::OutBuf ob = rr::out_buf(c, WS_ML_Y, y, batch * no);
    hipLaunchKernelGGL(k_mlp_store_rows, dim3(rr::grid_for(batch * no)), dim3(256), 0, c.stream, m->act.back(), batch,
                       (size_t)no, ldp(no), ob.dev);
    rr::finish_out(c, ob);
  });
}

// Loss and gradients without an update (gradient checks); grads may be NULL.
int rr_mlp_loss_grad(rr_handle h, rr_mlp_t m, size_t batch, const double* x, const double* target, double* loss,
                     float* grads) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(m && x && target && loss, RR_ERR_ARG, "mlp: null argument");
    RR_REQUIRE(batch > 0, RR_ERR_ARG, "mlp: empty batch");
    ensure_batch(m, batch);
    const double* dx = rr::in_buf(c, WS_ML_X, x, batch * m->dims.front());
    const double* dt = rr::in_buf(c, WS_ML_T, target, batch * m->dims.back());
    forward(c, m, batch, dx);
    backward(c, m, batch, dt);
    float lh = 0.f;
    RR_HIP(hipMemcpyAsync(&lh, m->loss, sizeof(float), hipMemcpyDeviceToHost, c.stream));
    if (grads) RR_HIP(hipMemcpyAsync(grads, m->grads, m->n_params * sizeof(float), hipMemcpyDefault, c.stream));
    RR_HIP(hipStreamSynchronize(c.stream));
    *loss = lh;
  });
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on a batch; returns the loss before the step.
int rr_mlp_train_step(rr_handle h, rr_mlp_t m, size_t batch, const double* x, const double* target, double lr,
                      double* loss) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(m && x && target, RR_ERR_ARG, "mlp: null argument");
    RR_REQUIRE(batch > 0, RR_ERR_ARG, "mlp: empty batch");
    ensure_batch(m, batch);
    const double* dx = rr::in_buf(c, WS_ML_X, x, batch * m->dims.front());
    const double* dt = rr::in_buf(c, WS_ML_T, target, batch * m->dims.back());
    forward(c, m, batch, dx);
    backward(c, m, batch, dt);
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(m->step)), c2 = 1.f - std::pow(b2, static_cast<float>(m->step));
    hipLaunchKernelGGL(k_mlp_adam, dim3(rr::grid_for(m->n_params)), dim3(256), 0, c.stream, m->params, m->grads, m->m1,
                       m->m2, m->n_params, static_cast<float>(lr), b1, b2, 1e-8f, c1, c2);
    RR_HIP(hipGetLastError());
    if (loss) {
      float lh = 0.f;
      RR_HIP(hipMemcpyAsync(&lh, m->loss, sizeof(float), hipMemcpyDeviceToHost, c.stream));
      RR_HIP(hipStreamSynchronize(c.stream));
      *loss = lh;
    }
  });
}

}  // extern "C"
