// librerdmft_gpu, batches of independent molecules: many small RI tensors of different sizes kept
// resident together, and one J/K build per molecule (factor form, D = L L^T) for the whole batch
// in a single HIP-graph launch.
//
// The first call for a given set of factor ranks and precision runs the batch eagerly (this also
// sizes the shared workspaces and lets the tile autotuner see every shape), then captures the same
// sequence of launches into a graph; later calls copy the new factors into the batch's staging
// buffer, replay the graph and copy the results out. Molecules run on up to eight parallel graph
// branches, each with a private stream and workspace set, so small molecules overlap. A graph is
// rebuilt whenever a workspace buffer it recorded has been reallocated since (a branch's workspace
// generation changed).
#include <algorithm>

#include "rr_core.hpp"

// One branch of the batch graph: a private stream and workspace set (kernel launches of different
// branches share nothing but the read-only tensors and disjoint output regions).
struct BatchLane {
  rr::Context ctx;
  hipEvent_t done = nullptr;
};

// One contiguous device-to-device copy of a segmented copy launch.
struct CopySeg {
  const double* src;
  double* dst;
  unsigned long long count;
  unsigned long long first_block;  // blocks of this segment: [first_block, next segment's first_block)
};

struct rr_batch {
  std::vector<rr::Tensor> t;
  std::vector<BatchLane*> lanes;
  std::vector<size_t> l_off, o_off;  // offsets (doubles) into the staging buffers
  std::vector<size_t> r;             // factor ranks of the captured graph
  int precision = -1;
  double* lin = nullptr;             // concatenated factors (device)
  double* jout = nullptr;            // concatenated J, K (device)
  double* kout = nullptr;
  size_t lin_cap = 0, out_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned long long> generation;  // lane workspace generations at capture
  hipEvent_t fork = nullptr;
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
  // the table is read by the kernel launched right after on the same stream; the synchronous copy
  // from pageable memory returns once the table is in device memory
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
}

constexpr size_t kMaxLanes = 8;

void free_lane(BatchLane* l) {
  if (!l) return;
  (void)hipStreamSynchronize(l->ctx.stream);
  for (auto& kv : l->ctx.ws) (void)hipFree(kv.second.first);
  if (l->ctx.zpage) (void)hipFree(l->ctx.zpage);
  if (l->done) (void)hipEventDestroy(l->done);
  if (l->ctx.stream) (void)hipStreamDestroy(l->ctx.stream);
  delete l;
}

BatchLane* make_lane(int device) {
  BatchLane* l = new BatchLane();
  l->ctx.device = device;
  if (hipStreamCreateWithFlags(&l->ctx.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&l->done, hipEventDisableTiming) != hipSuccess ||
      hipMalloc(&l->ctx.zpage, 256) != hipSuccess || hipMemset(l->ctx.zpage, 0, 256) != hipSuccess) {
    free_lane(l);
    throw Error{RR_ERR_HIP, "batch: cannot create a graph branch"};
  }
  return l;
}

// Molecules are dealt round-robin to the lanes, largest first, each lane in order on its stream.
// The lanes fork from and join back into the main stream through events (captured as graph edges).
void enqueue_all(Context& c, rr_batch* b) {
  RR_HIP(hipEventRecord(b->fork, c.stream));
  std::vector<size_t> order(b->t.size());
  for (size_t m = 0; m < order.size(); ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return b->t[x].n > b->t[y].n; });
  for (BatchLane* l : b->lanes) RR_HIP(hipStreamWaitEvent(l->ctx.stream, b->fork, 0));
  for (size_t i = 0; i < order.size(); ++i) {
    const size_t m = order[i];
    Context& lc = b->lanes[i % b->lanes.size()]->ctx;
    rr::enqueue_jk_factor(lc, b->t[m], b->precision, b->lin + b->l_off[m], b->r[m], b->jout + b->o_off[m],
                          b->kout + b->o_off[m]);
  }
  for (BatchLane* l : b->lanes) {
    RR_HIP(hipEventRecord(l->done, l->ctx.stream));
    RR_HIP(hipStreamWaitEvent(c.stream, l->done, 0));
  }
}

std::vector<unsigned long long> lane_generations(const rr_batch* b) {
  std::vector<unsigned long long> g;
  for (const BatchLane* l : b->lanes) g.push_back(l->ctx.ws_generation);
  return g;
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
    RR_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    const size_t nl = std::max<size_t>(1, std::min(kMaxLanes, n_mol));
    for (size_t i = 0; i < nl; ++i) b->lanes.push_back(make_lane(c.device));
  });
  if (rc != RR_OK) {
    for (BatchLane* l : b->lanes) free_lane(l);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto& t : b->t) rr::tensor_free(t);
    if (b->jout) (void)hipFree(b->jout);
    if (b->kout) (void)hipFree(b->kout);
    delete b;
    return rc;
  }
  *out = b;
  return RR_OK;
}

int rr_batch_destroy(rr_handle h, rr_batch_t b) {
  if (!h || !b) return RR_ERR_ARG;
  return guarded(h, [&](Context& c) {
    RR_HIP(hipStreamSynchronize(c.stream));
    drop_graph(b);
    for (BatchLane* l : b->lanes) free_lane(l);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto& t : b->t) rr::tensor_free(t);
    if (b->lin) (void)hipFree(b->lin);
    if (b->jout) (void)hipFree(b->jout);
    if (b->kout) (void)hipFree(b->kout);
    if (b->segs) (void)hipFree(b->segs);
    delete b;
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
    const bool same_shape = b->exec && b->precision == precision && b->r == std::vector<size_t>(r, r + nm) &&
                            b->generation == lane_generations(b);
    if (!same_shape) {
      drop_graph(b);
      if (ltot > b->lin_cap) {
        RR_HIP(hipStreamSynchronize(c.stream));
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
      enqueue_all(c, b);  // eager run: results of this call, workspace sizing, autotuning
      RR_HIP(hipStreamSynchronize(c.stream));
      const std::vector<unsigned long long> gen = lane_generations(b);
      const bool timing = rr::set_phase_timing(false);  // event timing synchronizes: not capturable
      RR_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
      try {
        enqueue_all(c, b);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(c.stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        rr::set_phase_timing(timing);
        throw;
      }
      rr::set_phase_timing(timing);
      RR_HIP(hipStreamEndCapture(c.stream, &b->graph));
      RR_REQUIRE(lane_generations(b) == gen, RR_ERR_INTERNAL, "batch: workspace reallocated during graph capture");
      RR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->generation = gen;
    }
    std::vector<CopySeg> outs;
    for (size_t m = 0; m < nm; ++m) {
      const size_t cnt = b->t[m].n * b->t[m].n;
      if (!cnt) continue;
      double* dst[2] = {j_out[m], k_out[m]};
      const double* src[2] = {b->jout + b->o_off[m], b->kout + b->o_off[m]};
      for (int q = 0; q < 2; ++q) {
        if (rr::is_device_ptr(dst[q]))
          outs.push_back({src[q], dst[q], cnt, 0});
        else
          RR_HIP(hipMemcpyAsync(dst[q], src[q], cnt * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      }
    }
    segmented_copy(c, b, outs);
    if (replayed) *replayed = rep;
  });
}

}  // extern "C"
