// Batches of stereo pairs of different sizes on one GPU (no counterpart in
// the reference, which processes one pair per call, main.py:49-70).
//
// Pairs are spread over a few HIP streams (longest first, each to the
// stream with the least work so far). Each stream owns one pipeline whose
// buffers are sized for the largest pair it runs; cbmv_pipeline_reshape
// points it at each pair in turn. Every pair's whole path (about 150 kernel
// launches) is captured once into a hipGraph and replayed, so a run of the
// batch is one graph launch per pair; pairs on different streams overlap,
// which fills the GPU with small pairs that could not fill it alone.
//
// The autotuner picks, per (W rounded up to 256, D) bucket, the fused
// aggregation's tile width (4, 8 or 16 columns per workgroup) by timing the
// post-processing stage of the bucket's first pair; results do not depend
// on the choice.
#include <algorithm>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "internal.hpp"

using namespace cbmv;

struct cbmv_batch {
  int n = 0;
  std::vector<int32_t> H, W, D;
  std::vector<int> stream_of;            // pair -> stream
  std::vector<std::vector<int>> queue;   // stream -> pairs, in launch order
  std::vector<hipStream_t> streams;
  std::vector<cbmv_pipeline*> pipes;     // one per stream
  std::vector<std::unique_ptr<DevBuf>> inL, inR, out;
  std::vector<int32_t> tile;             // per pair fused-aggregation tile (0 = default)
  std::vector<hipGraphExec_t> graphs;    // per pair, empty until captured
  std::vector<hipEvent_t> ev0, ev1;      // per pair start / end
  hipEvent_t evb = nullptr;              // batch start

  ~cbmv_batch() {
    for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
    for (hipGraphExec_t g : graphs)
      if (g) (void)hipGraphExecDestroy(g);
    for (hipEvent_t e : ev0)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev1)
      if (e) (void)hipEventDestroy(e);
    if (evb) (void)hipEventDestroy(evb);
    for (cbmv_pipeline* p : pipes) cbmv_pipeline_destroy(p);
    for (hipStream_t s : streams) (void)hipStreamDestroy(s);
  }
};

#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

static int drop_graphs(cbmv_batch* b) {
  for (hipGraphExec_t& g : b->graphs)
    if (g) {
      CBMV_HIP(hipGraphExecDestroy(g));
      g = nullptr;
    }
  return CBMV_OK;
}

extern "C" {

cbmv_batch* cbmv_batch_create(int32_t n, const int32_t* H, const int32_t* W, const int32_t* D,
                              const cbmv_params* params, int32_t n_streams) {
  clear_error();
  if (n < 1 || !H || !W || !D) {
    set_error("batch: need n >= 1 pairs and their sizes");
    return nullptr;
  }
  if (n_streams < 1 || n_streams > 16) {
    set_error("batch: n_streams must be in [1, 16] (got %d)", n_streams);
    return nullptr;
  }
  for (int i = 0; i < n; i++)
    if (H[i] < 16 || W[i] < 16 || D[i] < 2 || D[i] > 512) {
      set_error("batch: pair %d has size %dx%d, ndisp %d (need >= 16x16, 2 <= ndisp <= 512)", i,
                W[i], H[i], D[i]);
      return nullptr;
    }
  auto b = std::make_unique<cbmv_batch>();
  b->n = n;
  b->H.assign(H, H + n);
  b->W.assign(W, W + n);
  b->D.assign(D, D + n);
  const int ns = std::min<int>(n_streams, n);
  // longest pairs first, each to the least-loaded stream
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  auto vox = [&](int i) { return (int64_t)H[i] * W[i] * D[i]; };
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return vox(a) > vox(c); });
  std::vector<int64_t> load(ns, 0);
  b->stream_of.assign(n, 0);
  b->queue.assign(ns, {});
  for (int i : idx) {
    const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->stream_of[i] = s;
    load[s] += vox(i);
  }
  // launch order within a stream: pair index order
  for (int i = 0; i < n; i++) b->queue[b->stream_of[i]].push_back(i);
  for (int s = 0; s < ns; s++) {
    int cH = 16, cW = 16, cD = 2;
    for (int i : b->queue[s]) {
      cH = std::max(cH, H[i]);
      cW = std::max(cW, W[i]);
      cD = std::max(cD, D[i]);
    }
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      set_error("batch: hipStreamCreateWithFlags failed");
      return nullptr;
    }
    b->streams.push_back(st);
    cbmv_pipeline* pl = cbmv_pipeline_create(cH, cW, cD, params);
    if (!pl) return nullptr;  // error set by the pipeline
    b->pipes.push_back(pl);
    // pipelines on several streams share the device: one SGM sweep per path
    // (paired sweeps wait on each other across workgroups, which costs more
    // there than the traffic they save: 1.09 against 0.965 s for the 64-pair
    // batch on two streams)
    if (ns > 1 && cbmv_pipeline_set_sgm_pairs(pl, 0) != CBMV_OK) return nullptr;
  }
  b->tile.assign(n, params ? params->cbca_tile : 0);
  b->graphs.assign(n, nullptr);
  b->ev0.assign(n, nullptr);
  b->ev1.assign(n, nullptr);
  for (int i = 0; i < n; i++) {
    const size_t np = (size_t)H[i] * W[i];
    b->inL.push_back(std::make_unique<DevBuf>());
    b->inR.push_back(std::make_unique<DevBuf>());
    b->out.push_back(std::make_unique<DevBuf>());
    if (b->inL[i]->alloc(np) || b->inR[i]->alloc(np) || b->out[i]->alloc(np * 4)) return nullptr;
    if (hipEventCreate(&b->ev0[i]) != hipSuccess || hipEventCreate(&b->ev1[i]) != hipSuccess) {
      set_error("batch: hipEventCreate failed");
      return nullptr;
    }
  }
  if (hipEventCreate(&b->evb) != hipSuccess) {
    set_error("batch: hipEventCreate failed");
    return nullptr;
  }
  return b.release();
}

int cbmv_batch_set_forest(cbmv_batch* b, const cbmv_forest* f) {
  clear_error();
  CBMV_REQUIRE(b && f, "null argument");
  for (cbmv_pipeline* p : b->pipes) RC(cbmv_pipeline_set_forest(p, f));
  return drop_graphs(b);
}

int cbmv_batch_set_mlp(cbmv_batch* b, const float* mean, const float* inv_std, const float* w1,
                       const float* b1, const float* w2, const float* b2, const float* w3,
                       const float* b3, int32_t activation) {
  clear_error();
  CBMV_REQUIRE(b, "null batch");
  for (cbmv_pipeline* p : b->pipes)
    RC(cbmv_pipeline_set_mlp(p, mean, inv_std, w1, b1, w2, b2, w3, b3, activation));
  return drop_graphs(b);
}

int cbmv_batch_pair_buffers(cbmv_batch* b, int32_t i, uint8_t** d_left, uint8_t** d_right,
                            float** d_disp) {
  clear_error();
  CBMV_REQUIRE(b && d_left && d_right && d_disp, "null argument");
  CBMV_REQUIRE(i >= 0 && i < b->n, "pair %d out of range [0, %d)", i, b->n);
  *d_left = b->inL[i]->as<uint8_t>();
  *d_right = b->inR[i]->as<uint8_t>();
  *d_disp = b->out[i]->as<float>();
  return CBMV_OK;
}

int cbmv_batch_tune(cbmv_batch* b, int32_t reps, int32_t* tile_out) {
  clear_error();
  CBMV_REQUIRE(b && reps >= 1, "bad arguments");
  RC(drop_graphs(b));
  std::map<std::pair<int, int>, int> best;  // (W bucket, D) -> tile
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CBMV_HIP(hipEventCreate(&e0));
  CBMV_HIP(hipEventCreate(&e1));
  int rc = CBMV_OK;
  for (int i = 0; i < b->n && rc == CBMV_OK; i++) {
    const std::pair<int, int> key((b->W[i] + 255) / 256, b->D[i]);
    if (best.count(key)) continue;
    cbmv_pipeline* pl = b->pipes[b->stream_of[i]];
    hipStream_t s = b->streams[b->stream_of[i]];
    rc = cbmv_pipeline_reshape(pl, b->H[i], b->W[i], b->D[i]);
    if (rc) break;
    // one full run so that the post stage reads this pair's volumes
    rc = cbmv_pipeline_run(pl, b->inL[i]->as<uint8_t>(), b->inR[i]->as<uint8_t>(),
                           b->out[i]->as<float>(), s);
    float tbest = 0.f;
    int choice = 8;
    for (int t : {4, 8, 16}) {
      if (rc) break;
      rc = cbmv_pipeline_set_cbca_tile(pl, t);
      float ms = 0.f;
      for (int r = 0; r <= reps && !rc; r++) {  // the first repetition warms up
        if (r == 1 && hipEventRecord(e0, s) != hipSuccess) rc = CBMV_EHIP;
        if (!rc)
          rc = cbmv_pipeline_run_stage(pl, 3, b->inL[i]->as<uint8_t>(),
                                       b->inR[i]->as<uint8_t>(), b->out[i]->as<float>(), s);
      }
      if (rc) break;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        set_error("batch: timing failed");
        rc = CBMV_EHIP;
        break;
      }
      if (t == 4 || ms < tbest) {
        tbest = ms;
        choice = t;
      }
    }
    if (!rc) rc = cbmv_pipeline_set_cbca_tile(pl, 0);
    best[key] = choice;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  for (int i = 0; i < b->n; i++) {
    b->tile[i] = best[std::make_pair((b->W[i] + 255) / 256, b->D[i])];
    if (tile_out) tile_out[i] = b->tile[i];
  }
  return CBMV_OK;
}

int cbmv_batch_capture(cbmv_batch* b) {
  clear_error();
  CBMV_REQUIRE(b, "null batch");
  RC(drop_graphs(b));
  for (int i = 0; i < b->n; i++) {
    const int si = b->stream_of[i];
    cbmv_pipeline* pl = b->pipes[si];
    hipStream_t s = b->streams[si];
    RC(cbmv_pipeline_reshape(pl, b->H[i], b->W[i], b->D[i]));
    RC(cbmv_pipeline_set_cbca_tile(pl, b->tile[i]));
    CBMV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = cbmv_pipeline_run(pl, b->inL[i]->as<uint8_t>(), b->inR[i]->as<uint8_t>(),
                                     b->out[i]->as<float>(), s);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess || !g) {
      set_error("batch: stream capture of pair %d failed: %s", i, hipGetErrorString(ec));
      return CBMV_EHIP;
    }
    const hipError_t ei = hipGraphInstantiate(&b->graphs[i], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      set_error("batch: graph instantiation of pair %d failed: %s", i, hipGetErrorString(ei));
      b->graphs[i] = nullptr;
      return CBMV_EHIP;
    }
  }
  return CBMV_OK;
}

int cbmv_batch_run(cbmv_batch* b, float* exec_ms, float* done_ms, float* total_ms) {
  clear_error();
  CBMV_REQUIRE(b, "null batch");
  for (int i = 0; i < b->n; i++)
    CBMV_REQUIRE(b->graphs[i], "batch: pair %d has no graph (call cbmv_batch_capture)", i);
  hipStream_t s0 = b->streams[0];
  CBMV_HIP(hipEventRecord(b->evb, s0));
  for (size_t s = 1; s < b->streams.size(); s++)
    CBMV_HIP(hipStreamWaitEvent(b->streams[s], b->evb, 0));
  // interleave the streams' queues so that every stream starts at once
  size_t depth = 0;
  for (const auto& q : b->queue) depth = std::max(depth, q.size());
  for (size_t k = 0; k < depth; k++)
    for (size_t s = 0; s < b->streams.size(); s++) {
      if (k >= b->queue[s].size()) continue;
      const int i = b->queue[s][k];
      CBMV_HIP(hipEventRecord(b->ev0[i], b->streams[s]));
      CBMV_HIP(hipGraphLaunch(b->graphs[i], b->streams[s]));
      CBMV_HIP(hipEventRecord(b->ev1[i], b->streams[s]));
    }
  float tot = 0.f;
  for (int i = 0; i < b->n; i++) {
    CBMV_HIP(hipEventSynchronize(b->ev1[i]));
    float e = 0.f, d = 0.f;
    CBMV_HIP(hipEventElapsedTime(&e, b->ev0[i], b->ev1[i]));
    CBMV_HIP(hipEventElapsedTime(&d, b->evb, b->ev1[i]));
    if (exec_ms) exec_ms[i] = e;
    if (done_ms) done_ms[i] = d;
    tot = std::max(tot, d);
  }
  if (total_ms) *total_ms = tot;
  return CBMV_OK;
}

int32_t cbmv_batch_stream_of(const cbmv_batch* b, int32_t i) {
  if (!b || i < 0 || i >= b->n) return -1;
  return b->stream_of[i];
}

void cbmv_batch_destroy(cbmv_batch* b) { delete b; }

}  // extern "C"
