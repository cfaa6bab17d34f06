// librerdmft_gpu, batches of independent molecules: many small RI tensors of different sizes kept
// resident together, and one J/K build per molecule (factor form, D = L L^T) for the whole batch
// in a single HIP-graph launch.
//
// Every phase of the per-molecule build (factor packing, half-transform, gamma, rank-k update,
// Coulomb stream, reductions, assembly) runs as one grouped launch over all molecules (rr_lib.hip:
// batch_plan_*), each molecule with the grid its single-molecule launch would use, so results equal
// rr_ri_jk_factor bitwise. The first call for a given set of factor ranks and precision builds the
// plan (segment tables and work buffers), runs it eagerly and captures its launches into a graph;
// later calls copy the new factors into the batch's staging buffer, replay the graph and copy the
// results out, each copy direction as one segmented copy launch.
#include <algorithm>

#include "rr_core.hpp"

// One contiguous device-to-device copy of a segmented copy launch.
struct CopySeg {
  const double* src;
  double* dst;
  unsigned long long count;
  unsigned long long first_block;  // blocks of this segment: [first_block, next segment's first_block)
};

struct rr_batch {
  std::vector<rr::Tensor> t;
  std::vector<size_t> l_off, o_off;  // offsets (doubles) into the staging buffers
  std::vector<size_t> r;             // factor ranks of the captured graph
  int precision = -1;
  double* lin = nullptr;             // concatenated factors (device)
  double* jout = nullptr;            // concatenated J, K (device)
  double* kout = nullptr;
  size_t lin_cap = 0, out_cap = 0;
  rr::BatchPlan* plan = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  CopySeg* segs = nullptr;           // device table of the per-call gather / scatter copies
  size_t seg_cap = 0;
};

// Many device-to-device copies in one launch: block b copies a 4 KiB chunk of the segment whose
// block range holds b (found by binary search over the table's first_block column).
__global__ __launch_bounds__(256) void k_segment_copy(const CopySeg* segs, int nseg) {
  int lo = 0, hi = nseg;
  const unsigned long long b = blockIdx.x;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (segs[mid].first_block <= b) lo = mid; else hi = mid;
  }
  const CopySeg sg = segs[lo];
  const unsigned long long base = (b - sg.first_block) * 512;
  for (unsigned long long i = base + threadIdx.x; i < base + 512 && i < sg.count; i += blockDim.x) sg.dst[i] = sg.src[i];
}

namespace {

using rr::Context;
using rr::Error;

constexpr unsigned long long kCopyChunk = 512;  // doubles per block of k_segment_copy

// Device-to-device copies (src, dst, count) in one launch; returns false if there is nothing to copy.
bool segmented_copy(Context& c, rr_batch* b, const std::vector<CopySeg>& in) {
  std::vector<CopySeg> segs;
  unsigned long long blocks = 0;
  for (const CopySeg& sg : in) {
    if (!sg.count) continue;
    CopySeg x = sg;
    x.first_block = blocks;
    blocks += (sg.count + kCopyChunk - 1) / kCopyChunk;
    segs.push_back(x);
  }
  if (segs.empty()) return false;
  RR_REQUIRE(blocks < (1ull << 31), RR_ERR_ARG, "batch: copy too large");
  if (segs.size() > b->seg_cap) {
    RR_HIP(hipStreamSynchronize(c.stream));
    if (b->segs) RR_HIP(hipFree(b->segs));
    b->segs = nullptr;
    RR_HIP(hipMalloc(&b->segs, segs.size() * sizeof(CopySeg)));
    b->seg_cap = segs.size();
  }
  // the table is read by the kernel launched right after on the same stream; the host vector is
  // released on return, so wait for the copy
  RR_HIP(hipMemcpyAsync(b->segs, segs.data(), segs.size() * sizeof(CopySeg), hipMemcpyHostToDevice, c.stream));
  RR_HIP(hipStreamSynchronize(c.stream));
  hipLaunchKernelGGL(k_segment_copy, dim3((unsigned)blocks), dim3(256), 0, c.stream, b->segs, (int)segs.size());
  RR_HIP(hipGetLastError());
  return true;
}

void drop_graph(rr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  rr::batch_plan_free(b->plan);
  b->plan = nullptr;
}

void free_batch(rr_batch* b) {
  drop_graph(b);
  for (auto& t : b->t) rr::tensor_free(t);
  if (b->lin) (void)hipFree(b->lin);
  if (b->jout) (void)hipFree(b->jout);
  if (b->kout) (void)hipFree(b->kout);
  if (b->segs) (void)hipFree(b->segs);
  delete b;
}

}  // namespace

using rr::guarded;

extern "C" {

int rr_batch_create(rr_handle h, size_t n_mol, const size_t* n_aux, const size_t* n, const double* const* eri3,
                    unsigned copies, rr_batch_t* out) {
  if (!out) return RR_ERR_ARG;
  *out = nullptr;
  rr_batch* b = new rr_batch();
  const int rc = guarded(h, [&](Context& c) {
    RR_REQUIRE(n_mol == 0 || (n_aux && n && eri3), RR_ERR_ARG, "batch: null arrays");
    b->t.resize(n_mol);
    size_t off = 0;
    b->o_off.resize(n_mol);
    for (size_t m = 0; m < n_mol; ++m) {
      rr::tensor_store(c, b->t[m], eri3[m], n_aux[m], n[m], copies);
      b->o_off[m] = off;
      off += n[m] * n[m];
    }
    b->out_cap = std::max<size_t>(off, 1);
    RR_HIP(hipMalloc(&b->jout, b->out_cap * sizeof(double)));
    RR_HIP(hipMalloc(&b->kout, b->out_cap * sizeof(double)));
  });
  if (rc != RR_OK) {
    free_batch(b);
    return rc;
  }
  *out = b;
  return RR_OK;
}

int rr_batch_destroy(rr_handle h, rr_batch_t b) {
  if (!h || !b) return RR_ERR_ARG;
  return guarded(h, [&](Context& c) {
    RR_HIP(hipStreamSynchronize(c.stream));
    free_batch(b);
  });
}

int rr_batch_jk_factor(rr_handle h, rr_batch_t b, int precision, const size_t* r, const double* const* lfac,
                       double* const* j_out, double* const* k_out, int* replayed) {
  return guarded(h, [&](Context& c) {
    RR_REQUIRE(b, RR_ERR_ARG, "batch: null batch");
    rr::check_prec(precision);
    const size_t nm = b->t.size();
    RR_REQUIRE(nm == 0 || (r && lfac && j_out && k_out), RR_ERR_ARG, "batch: null arrays");
    for (size_t m = 0; m < nm; ++m) {
      const rr::Tensor& t = b->t[m];
      RR_REQUIRE(t.f64 || t.f32 || t.b16, RR_ERR_STATE, "batch: empty tensor");
      (void)rr::tensor_data(t, precision);  // the requested copy exists
      RR_REQUIRE(r[m] == 0 || lfac[m], RR_ERR_ARG, "batch: null factor");
    }
    // staging layout for these ranks
    std::vector<size_t> loff(nm);
    size_t ltot = 0;
    for (size_t m = 0; m < nm; ++m) {
      loff[m] = ltot;
      ltot += b->t[m].n * r[m];
    }
    const bool same_shape = b->exec && b->precision == precision && b->r == std::vector<size_t>(r, r + nm);
    if (!same_shape) {
      RR_HIP(hipStreamSynchronize(c.stream));
      drop_graph(b);
      if (ltot > b->lin_cap) {
        if (b->lin) RR_HIP(hipFree(b->lin));
        b->lin = nullptr;
        RR_HIP(hipMalloc(&b->lin, std::max<size_t>(ltot, 1) * sizeof(double)));
        b->lin_cap = ltot;
      }
      b->l_off = loff;
      b->r.assign(r, r + nm);
      b->precision = precision;
    }
    // factors into the staging buffer (host arrays are gathered into one transfer)
    std::vector<double> host;
    std::vector<CopySeg> dcopy;
    for (size_t m = 0; m < nm; ++m) {
      const size_t cnt = b->t[m].n * r[m];
      if (!cnt) continue;
      if (rr::is_device_ptr(lfac[m])) {
        RR_VALIDATE_EXTENT(lfac[m], cnt * sizeof(double), "batch factor");
        dcopy.push_back({lfac[m], b->lin + loff[m], cnt, 0});
      } else {
        if (host.empty()) host.assign(ltot, 0.0);
        std::copy(lfac[m], lfac[m] + cnt, host.begin() + loff[m]);
      }
    }
    segmented_copy(c, b, dcopy);
    if (!host.empty()) {
      // device-pointer factors were copied in place; copy the host part around them
      for (size_t m = 0; m < nm; ++m) {
        const size_t cnt = b->t[m].n * r[m];
        if (cnt && !rr::is_device_ptr(lfac[m]))
          RR_HIP(hipMemcpyAsync(b->lin + loff[m], host.data() + loff[m], cnt * sizeof(double), hipMemcpyHostToDevice,
                                c.stream));
      }
      RR_HIP(hipStreamSynchronize(c.stream));  // `host` is released at the end of this scope
    }
    int rep = 0;
    if (same_shape) {
      RR_HIP(hipGraphLaunch(b->exec, c.stream));
      rep = 1;
    } else {
      b->plan = rr::batch_plan_create(c, b->t, precision, b->r, b->lin, b->l_off, b->jout, b->kout, b->o_off);
      rr::batch_plan_enqueue(c, b->plan);  // eager run: the results of this call
      RR_HIP(hipStreamSynchronize(c.stream));
      RR_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
      try {
        rr::batch_plan_enqueue(c, b->plan);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(c.stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      RR_HIP(hipStreamEndCapture(c.stream, &b->graph));
      RR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    std::vector<CopySeg> outs;
    for (size_t m = 0; m < nm; ++m) {
      const size_t cnt = b->t[m].n * b->t[m].n;
      if (!cnt) continue;
      double* dst[2] = {j_out[m], k_out[m]};
      const double* src[2] = {b->jout + b->o_off[m], b->kout + b->o_off[m]};
      for (int q = 0; q < 2; ++q) {
        if (rr::is_device_ptr(dst[q])) {
          RR_VALIDATE_EXTENT(dst[q], cnt * sizeof(double), "batch output");
          outs.push_back({src[q], dst[q], cnt, 0});
        } else
          RR_HIP(hipMemcpyAsync(dst[q], src[q], cnt * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      }
    }
    segmented_copy(c, b, outs);
    if (replayed) *replayed = rep;
  });
}

}  // extern "C"
