// batch.cpp -- many scenes checked together on one device.
//
// No reference counterpart: the reference holds one scene per
// RobotConfigspace object and launches one kernel per request
// (robotconfigspace.cpp:718-723).  A batch takes initialised contexts (one
// scene each, same device, same ndof) and runs the point checks of all of
// them as one sequence: a fill of the work counters, ONE kinematics launch
// over every scene's configurations, one persistent separating-vector launch
// per vertex-count bucket (every scene of the bucket, blocks rotating over
// the scenes, see sva_multi_kernel), and ONE finalize launch.  With use_graph
// the sequence is captured once into a hipGraph and replayed for as long as
// the call's arrays and sizes stay the same.
//
// Buckets group scenes by total vertex count; each bucket has its own launch
// shape (block size and occupancy target, i.e. how many scene copies share
// a CU's LDS, and pairs per work unit), set by the caller or picked by
// gprm_batch_autotune from timed runs of the caller's workload.
//
// Edge checks of many scenes (gprm_batch_indicator2_dev) follow the
// single-scene edge pipeline (run_edges in gprm_api.cpp) with every stage a
// multi-scene launch: ONE prologue over all scenes' edges (testnum, sample
// order, largest testnum per scene, read back once), then launch steps, each
// ONE compaction and ONE kinematics launch over every scene's part of the
// step and one persistent separating-vector launch per bucket.  Steps walk
// the order-position windows [0,12), [12,60), ... shared by all scenes; a
// window whose items exceed a scene's per-step budget takes several steps,
// each over a range of that scene's edges.  The steps after the prologue
// are what a hipGraph captures; the graph is kept while the arrays, sizes
// and the per-scene largest testnums stay the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gprm.h"
#include "context.h"
#include "runtime.h"

namespace gprm {
namespace {

// Bucket edges in total scene vertices: [0,512), [512,1024), [1024,2048), [2048, inf).
constexpr int kBucketEdges[] = {0, 512, 1024, 2048, 1 << 30};
constexpr int kBuckets = 4;
constexpr int kMaxBatchItems = 1 << 22;  // configurations per scene and call

struct Bucket {
  int lo = 0, hi = 0;          // vertex-count range
  std::vector<int> scenes;     // indices into the batch
  size_t lds = 0;              // largest scene image among its scenes (bytes)
  SvaShape shape;
  DevBuf tasks;                // SceneTask[scenes.size()]
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct Signature {
  std::vector<const void*> q, res;
  std::vector<long long> off;
  std::vector<int> n;
  bool operator==(const Signature& o) const { return q == o.q && res == o.res && off == o.off && n == o.n; }
};

constexpr int kMaxBatchEdges = 1 << 22;   // edges per scene and call
constexpr int kEdgeStepItems = 1 << 19;   // samples per scene and launch step

struct EdgeSignature {
  std::vector<const void*> qs, qe, res;
  std::vector<long long> off;
  std::vector<int> n;
  bool operator==(const EdgeSignature& o) const {
    return qs == o.qs && qe == o.qe && res == o.res && off == o.off && n == o.n;
  }
};

// One launch step of a multi-scene edge call.
struct EdgeStep {
  std::vector<EdgeTask> tasks;   // one per scene with a part in the step, by scene
  std::vector<int> scene;        // the scene of each task
  int cblocks = 0, fblocks = 0;  // compaction and kinematics grids
  DevBuf d_tasks;                // EdgeTask[tasks.size()]
  DevBuf sva[4];                 // per bucket: SceneTask of the bucket's tasks
  int nsva[4] = {0, 0, 0, 0};
};

}  // namespace
}  // namespace gprm

using namespace gprm;

struct gprm_batch {
  std::vector<gprm_ctx*> ctx;
  int device = 0, ndof = 0;
  hipStream_t stream = nullptr;
  Bucket buckets[kBuckets];
  DevBuf frames, words, ctl, fk_tasks, hist;
  int fk_blocks = 0;
  long long total_items = 0;
  Signature sig;
  bool have_sig = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int clock_khz = 0;
  std::string err;
  // multi-scene edge calls
  EdgeSignature esig;
  bool have_esig = false;
  std::vector<long long> e_base;    // per scene: first edge in e_testnum / e_stride
  DevBuf e_testnum, e_stride, e_max, e_total, e_clock, e_pro;
  int e_pro_blocks = 0, e_pro_tasks = 0;
  std::vector<int> e_plan_key;      // per scene largest testnum of the current plan
  bool have_eplan = false;
  std::vector<std::unique_ptr<EdgeStep>> e_steps;
  DevBuf e_items, e_frames, e_ctl;  // step buffers; e_ctl: (items, work counter, units done) per task
  int e_ntasks = 0;
  hipGraph_t egraph = nullptr;
  hipGraphExec_t eexec = nullptr;
  ~gprm_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (eexec) (void)hipGraphExecDestroy(eexec);
    if (egraph) (void)hipGraphDestroy(egraph);
    for (auto& b : buckets) {
      if (b.ev0) (void)hipEventDestroy(b.ev0);
      if (b.ev1) (void)hipEventDestroy(b.ev1);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

thread_local std::string g_batch_error;

int bfail(gprm_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  else g_batch_error = msg;
  return code;
}

int bhip(gprm_batch* b, hipError_t e, const char* what) {
  return bfail(b, GPRM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define BHIP(b, expr, what)                              \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return bhip(b, e_, what);      \
  } while (0)

void drop_edge_plan(gprm_batch* b) {
  if (b->eexec) (void)hipGraphExecDestroy(b->eexec);
  if (b->egraph) (void)hipGraphDestroy(b->egraph);
  b->eexec = nullptr;
  b->egraph = nullptr;
  b->have_eplan = false;
}

void drop_graph(gprm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->have_sig = false;
  drop_edge_plan(b);  // its separating-vector tables carry the bucket shapes too
}

// Control words: per scene (work counter, units done), then per scene the
// completion clock (u64), then the start clock (u64).
size_t ctl_words(int S) { return 2 * (size_t)S + 2 * (size_t)S + 2; }
unsigned* ctl_counter(gprm_batch* b, int s) { return b->ctl.as<unsigned>() + 2 * s; }
unsigned long long* ctl_tdone(gprm_batch* b, int s) {
  return reinterpret_cast<unsigned long long*>(b->ctl.as<unsigned>() + 2 * b->ctx.size()) + s;
}
unsigned long long* ctl_tstart(gprm_batch* b) {
  return reinterpret_cast<unsigned long long*>(b->ctl.as<unsigned>() + 4 * b->ctx.size());
}

// Task tables and buffers for one call signature (uploaded outside any capture).
int prepare(gprm_batch* b, const Signature& sg) {
  const int S = (int)b->ctx.size();
  const int nd = b->ndof;
  std::vector<long long> base(S + 1, 0);
  for (int s = 0; s < S; ++s) base[s + 1] = base[s] + frame_stride(sg.n[s]);
  b->total_items = 0;
  for (int s = 0; s < S; ++s) b->total_items += sg.n[s];
  BHIP(b, b->frames.ensure(sizeof(float) * 12 * nd * (size_t)std::max(1LL, base[S])), "hipMalloc frames");
  BHIP(b, b->words.ensure(sizeof(int) * (size_t)std::max(1LL, base[S])), "hipMalloc words");
  BHIP(b, b->ctl.ensure(sizeof(unsigned) * ctl_words(S)), "hipMalloc control words");
  size_t hist = 0;
  for (auto& bk : b->buckets)
    if (!bk.scenes.empty()) hist = std::max(hist, (size_t)kMaxIter * bk.shape.grid * bk.shape.block);
  const void* old_hist = b->hist.p;
  BHIP(b, b->hist.ensure(sizeof(float) * 4 * std::max<size_t>(hist, 1)), "hipMalloc history");
  if (b->hist.p != old_hist) drop_edge_plan(b);  // the edge steps' tables hold the old address
  std::vector<FkTask> fk(S);
  int blocks = 0;
  for (int s = 0; s < S; ++s) {
    FkTask& t = fk[s];
    t.rb = b->ctx[s]->rb;
    t.q = static_cast<const float*>(sg.q[s]);
    t.offset = sg.off[s];
    t.n = sg.n[s];
    t.block0 = blocks;
    t.fstride = frame_stride(sg.n[s]);
    t.frames = b->frames.as<float>() + (size_t)12 * nd * base[s];
    t.word = b->words.as<int>() + base[s];
    t.codes = static_cast<int*>(const_cast<void*>(sg.res[s]));
    blocks += (sg.n[s] + kFkBlock - 1) / kFkBlock;
  }
  b->fk_blocks = blocks;
  BHIP(b, b->fk_tasks.ensure(sizeof(FkTask) * (size_t)S), "hipMalloc tasks");
  BHIP(b, hipMemcpy(b->fk_tasks.p, fk.data(), sizeof(FkTask) * (size_t)S, hipMemcpyHostToDevice), "copy tasks");
  for (auto& bk : b->buckets) {
    if (bk.scenes.empty()) continue;
    std::vector<SceneTask> st(bk.scenes.size());
    for (size_t k = 0; k < bk.scenes.size(); ++k) {
      const int s = bk.scenes[k];
      const gprm_ctx* c = b->ctx[s];
      SceneTask& t = st[k];
      t.a = SvaArgs{};
      t.a.img = c->img;
      t.a.ndof = nd;
      t.a.frames = fk[s].frames;
      t.a.fstride = fk[s].fstride;
      t.a.hist = b->hist.as<float4>();
      t.a.nitems = sg.n[s];
      t.a.counter = ctl_counter(b, s);
      t.a.pairs_per_unit = bk.shape.ppu;
      t.a.units_per_item = (c->img.M + bk.shape.ppu - 1) / bk.shape.ppu;
      if ((long long)sg.n[s] * t.a.units_per_item >= (1LL << 31))
        return bfail(b, GPRM_E_CAPACITY, "too many work units for scene " + std::to_string(s));
      t.a.word = fk[s].word;
      t.done = ctl_counter(b, s) + 1;
      t.t_done = ctl_tdone(b, s);
    }
    BHIP(b, bk.tasks.ensure(sizeof(SceneTask) * st.size()), "hipMalloc tasks");
    BHIP(b, hipMemcpy(bk.tasks.p, st.data(), sizeof(SceneTask) * st.size(), hipMemcpyHostToDevice), "copy tasks");
  }
  return GPRM_OK;
}

// The call's launch sequence on b->stream (capturable: kernels only).
int enqueue(gprm_batch* b, bool events) {
  const int S = (int)b->ctx.size();
  hipStream_t st = b->stream;
  BHIP(b, launch_fill_u32(b->ctl.as<unsigned>(), 0u, 4LL * S, st), "fill");
  BHIP(b, launch_fill_u32(b->ctl.as<unsigned>() + 4 * S, 0xffffffffu, 2, st), "fill");
  BHIP(b, launch_fk_points_multi(b->fk_tasks.as<FkTask>(), S, b->fk_blocks, b->ndof, st), "kinematics launch");
  for (auto& bk : b->buckets) {
    if (bk.scenes.empty()) continue;
    if (events) BHIP(b, hipEventRecord(bk.ev0, st), "event");
    BHIP(b, launch_sva_multi(bk.tasks.as<SceneTask>(), (int)bk.scenes.size(), 0, bk.shape, bk.lds,
                             ctl_tstart(b), nullptr, st),
         "separating-vector launch");
    if (events) BHIP(b, hipEventRecord(bk.ev1, st), "event");
  }
  BHIP(b, launch_finalize_points_multi(b->fk_tasks.as<FkTask>(), S, b->fk_blocks, st), "finalize launch");
  return GPRM_OK;
}

int run(gprm_batch* b, const Signature& sg, bool use_graph, bool events, float* done_ms) {
  BHIP(b, hipSetDevice(b->device), "hipSetDevice");
  const bool same = b->have_sig && b->sig == sg;
  if (!same) {
    drop_graph(b);
    int rc = prepare(b, sg);
    if (rc != GPRM_OK) return rc;
    b->sig = sg;
    b->have_sig = true;
  }
  if (use_graph && !events) {
    if (!b->exec) {
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      int rc = enqueue(b, false);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc != GPRM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return bhip(b, e, "end capture");
      b->graph = g;
      BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate graph");
    }
    BHIP(b, hipGraphLaunch(b->exec, b->stream), "graph launch");
  } else {
    int rc = enqueue(b, events);
    if (rc != GPRM_OK) return rc;
  }
  BHIP(b, hipStreamSynchronize(b->stream), "sync");
  if (done_ms) {
    const int S = (int)b->ctx.size();
    std::vector<unsigned long long> td(S + 1);
    BHIP(b, hipMemcpy(td.data(), ctl_tdone(b, 0), sizeof(unsigned long long) * (S + 1), hipMemcpyDeviceToHost),
         "copy clocks");
    const unsigned long long t0 = td[S];
    for (int s = 0; s < S; ++s)
      done_ms[s] = (sg.n[s] == 0 || td[s] == 0 || td[s] < t0) ? 0.0f
                                                               : (float)((double)(td[s] - t0) / b->clock_khz);
  }
  return GPRM_OK;
}

int signature(gprm_batch* b, const float* const* d_q, const long long* offset, const int* n, int* const* d_res,
              Signature* sg) {
  const int S = (int)b->ctx.size();
  if (!d_q || !offset || !n || !d_res) return bfail(b, GPRM_E_INVALID, "null argument");
  sg->q.resize(S);
  sg->res.resize(S);
  sg->off.resize(S);
  sg->n.resize(S);
  for (int s = 0; s < S; ++s) {
    if (n[s] < 0 || n[s] > kMaxBatchItems || offset[s] < n[s] || (n[s] > 0 && (!d_q[s] || !d_res[s])))
      return bfail(b, GPRM_E_INVALID, "bad arguments for scene " + std::to_string(s));
    sg->q[s] = d_q[s];
    sg->res[s] = d_res[s];
    sg->off[s] = offset[s];
    sg->n[s] = n[s];
  }
  return GPRM_OK;
}

// ---- multi-scene edge calls -------------------------------------------------

int edge_signature(gprm_batch* b, const float* const* d_qs, const float* const* d_qe, const long long* offset,
                   const int* n, int* const* d_res, EdgeSignature* sg) {
  const int S = (int)b->ctx.size();
  if (!d_qs || !d_qe || !offset || !n || !d_res) return bfail(b, GPRM_E_INVALID, "null argument");
  sg->qs.resize(S);
  sg->qe.resize(S);
  sg->res.resize(S);
  sg->off.resize(S);
  sg->n.resize(S);
  for (int s = 0; s < S; ++s) {
    if (n[s] < 0 || n[s] > kMaxBatchEdges || offset[s] < n[s] ||
        (n[s] > 0 && (!d_qs[s] || !d_qe[s] || !d_res[s])))
      return bfail(b, GPRM_E_INVALID, "bad arguments for scene " + std::to_string(s));
    sg->qs[s] = d_qs[s];
    sg->qe[s] = d_qe[s];
    sg->res[s] = d_res[s];
    sg->off[s] = offset[s];
    sg->n[s] = n[s];
  }
  return GPRM_OK;
}

EdgeArgs edge_args(gprm_batch* b, const EdgeSignature& sg, int s) {
  EdgeArgs ea{};
  ea.qs = static_cast<const float*>(sg.qs[s]);
  ea.qe = static_cast<const float*>(sg.qe[s]);
  ea.offset = sg.off[s];
  ea.nedges = sg.n[s];
  ea.testnum = b->e_testnum.as<int>() + b->e_base[s];
  ea.stride = b->e_stride.as<int>() + b->e_base[s];
  ea.edge_res = static_cast<int*>(const_cast<void*>(sg.res[s]));
  return ea;
}

// Per-signature buffers and the prologue task table (outside any capture).
int edge_prepare(gprm_batch* b, const EdgeSignature& sg) {
  const int S = (int)b->ctx.size();
  b->e_base.assign(S + 1, 0);
  for (int s = 0; s < S; ++s) b->e_base[s + 1] = b->e_base[s] + sg.n[s];
  const size_t E = (size_t)std::max(1LL, b->e_base[S]);
  BHIP(b, b->e_testnum.ensure(sizeof(int) * E), "hipMalloc testnum");
  BHIP(b, b->e_stride.ensure(sizeof(int) * E), "hipMalloc stride");
  BHIP(b, b->e_max.ensure(sizeof(int) * (size_t)S), "hipMalloc");
  BHIP(b, b->e_total.ensure(sizeof(unsigned long long) * (size_t)S), "hipMalloc");
  // completion clocks per scene, then the start clock
  BHIP(b, b->e_clock.ensure(sizeof(unsigned long long) * (size_t)(S + 1)), "hipMalloc");
  std::vector<EdgeTask> pro;
  int blocks = 0;
  for (int s = 0; s < S; ++s) {
    if (sg.n[s] == 0) continue;
    EdgeTask t{};
    t.ea = edge_args(b, sg, s);
    t.rb = b->ctx[s]->rb;
    t.dq = b->ctx[s]->dq;
    t.testnum = b->e_testnum.as<int>() + b->e_base[s];
    t.stride = b->e_stride.as<int>() + b->e_base[s];
    t.max_testnum = b->e_max.as<int>() + s;
    t.total = b->e_total.as<unsigned long long>() + s;
    t.block0 = blocks;
    blocks += (sg.n[s] + 255) / 256;
    pro.push_back(t);
  }
  b->e_pro_blocks = blocks;
  b->e_pro_tasks = (int)pro.size();
  if (!pro.empty()) {
    BHIP(b, b->e_pro.ensure(sizeof(EdgeTask) * pro.size()), "hipMalloc tasks");
    BHIP(b, hipMemcpy(b->e_pro.p, pro.data(), sizeof(EdgeTask) * pro.size(), hipMemcpyHostToDevice),
         "copy tasks");
  }
  return GPRM_OK;
}

// Launch steps for the per-scene largest testnums maxT (outside any capture).
int edge_plan(gprm_batch* b, const EdgeSignature& sg, const std::vector<int>& maxT) {
  const int S = (int)b->ctx.size();
  const int nd = b->ndof;
  int gmax = 0;
  for (int s = 0; s < S; ++s) gmax = std::max(gmax, sg.n[s] > 0 ? maxT[s] : 0);
  // per scene: item budget of a step (the unit count must stay below 2^31
  // for the smallest units, one pair each) and the capacity it needs
  std::vector<int> budget(S), cap(S, 0);
  for (int s = 0; s < S; ++s)
    budget[s] = std::max(1, std::min(kEdgeStepItems, 0x7fffffff / std::max(1, b->ctx[s]->img.M)));
  struct Part { int s, e0, e1, lo, hi, cap; };
  std::vector<std::vector<Part>> steps;
  for (const auto& w : edge_windows(gmax)) {
    std::vector<int> per(S, 0), nch(S, 0);
    int nsteps = 0;
    for (int s = 0; s < S; ++s) {
      if (sg.n[s] == 0 || maxT[s] <= w.first) continue;
      const int width = std::min(w.second, maxT[s]) - w.first;
      per[s] = std::max(1, budget[s] / width);
      nch[s] = (sg.n[s] + per[s] - 1) / per[s];
      nsteps = std::max(nsteps, nch[s]);
    }
    for (int j = 0; j < nsteps; ++j) {
      std::vector<Part> parts;
      for (int s = 0; s < S; ++s) {
        if (j >= nch[s]) continue;
        const int width = std::min(w.second, maxT[s]) - w.first;
        const int e0 = j * per[s], e1 = std::min(sg.n[s], e0 + per[s]);
        const long long c = frame_stride((long long)(e1 - e0) * width);
        parts.push_back({s, e0, e1, w.first, w.second, (int)c});
        cap[s] = std::max(cap[s], (int)c);
      }
      steps.push_back(std::move(parts));
    }
  }
  // step buffers: items and frames per scene (regions sized by its largest
  // step), control words per task
  std::vector<long long> ibase(S + 1, 0);
  for (int s = 0; s < S; ++s) ibase[s + 1] = ibase[s] + cap[s];
  int ntasks = 0;
  for (const auto& st : steps) ntasks += (int)st.size();
  BHIP(b, b->e_items.ensure(sizeof(int) * 2 * (size_t)std::max(1LL, ibase[S])), "hipMalloc items");
  BHIP(b, b->e_frames.ensure(sizeof(float) * 12 * nd * (size_t)std::max(1LL, ibase[S])), "hipMalloc frames");
  BHIP(b, b->e_ctl.ensure(sizeof(int) * 3 * (size_t)std::max(1, ntasks)), "hipMalloc control words");
  size_t hist = 0;
  for (auto& bk : b->buckets)
    if (!bk.scenes.empty()) hist = std::max(hist, (size_t)kMaxIter * bk.shape.grid * bk.shape.block);
  const void* old_hist = b->hist.p;
  BHIP(b, b->hist.ensure(sizeof(float) * 4 * std::max<size_t>(hist, 1)), "hipMalloc history");
  if (b->hist.p != old_hist) {  // the point path's task tables hold the old address
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->have_sig = false;
  }
  int* const items = b->e_items.as<int>();
  const long long icap = std::max(1LL, ibase[S]);
  std::vector<int> bucket_of(S, -1);
  for (int i = 0; i < kBuckets; ++i)
    for (int s : b->buckets[i].scenes) bucket_of[s] = i;
  b->e_steps.clear();
  int task = 0;
  for (const auto& parts : steps) {
    auto step = std::make_unique<EdgeStep>();
    int cb = 0, fb = 0;
    std::vector<SceneTask> sva[kBuckets];
    for (const Part& pt : parts) {
      const int s = pt.s;
      const gprm_ctx* c = b->ctx[s];
      int* ctl = b->e_ctl.as<int>() + 3 * task++;
      EdgeTask t{};
      t.ea = edge_args(b, sg, s);
      t.rb = c->rb;
      t.dq = c->dq;
      t.e0 = pt.e0;
      t.e1 = pt.e1;
      t.lo = pt.lo;
      t.hi = pt.hi;
      t.item_edge = items + ibase[s];
      t.item_pos = items + icap + ibase[s];
      t.nitems = ctl;
      t.cap = pt.cap;
      t.frames = b->e_frames.as<float>() + (size_t)12 * nd * ibase[s];
      t.block0 = cb;
      t.fblock0 = fb;
      cb += (pt.e1 - pt.e0 + 255) / 256;
      fb += (pt.cap + kFkBlock - 1) / kFkBlock;
      step->tasks.push_back(t);
      step->scene.push_back(s);
      const Bucket& bk = b->buckets[bucket_of[s]];
      SceneTask q{};
      q.a.img = c->img;
      q.a.ndof = nd;
      q.a.frames = t.frames;
      q.a.fstride = pt.cap;
      q.a.hist = b->hist.as<float4>();
      q.a.nitems = -1;
      q.a.nitems_dev = ctl;
      q.a.counter = reinterpret_cast<unsigned*>(ctl + 1);
      q.a.pairs_per_unit = bk.shape.ppu;
      q.a.units_per_item = (c->img.M + bk.shape.ppu - 1) / bk.shape.ppu;
      q.a.item_edge = t.item_edge;
      q.a.edge_res = t.ea.edge_res;
      q.done = reinterpret_cast<unsigned*>(ctl + 2);
      q.t_done = b->e_clock.as<unsigned long long>() + s;
      sva[bucket_of[s]].push_back(q);
    }
    step->cblocks = cb;
    step->fblocks = fb;
    BHIP(b, step->d_tasks.ensure(sizeof(EdgeTask) * step->tasks.size()), "hipMalloc tasks");
    BHIP(b, hipMemcpy(step->d_tasks.p, step->tasks.data(), sizeof(EdgeTask) * step->tasks.size(),
                      hipMemcpyHostToDevice),
         "copy tasks");
    for (int i = 0; i < kBuckets; ++i) {
      step->nsva[i] = (int)sva[i].size();
      if (sva[i].empty()) continue;
      BHIP(b, step->sva[i].ensure(sizeof(SceneTask) * sva[i].size()), "hipMalloc tasks");
      BHIP(b, hipMemcpy(step->sva[i].p, sva[i].data(), sizeof(SceneTask) * sva[i].size(), hipMemcpyHostToDevice),
           "copy tasks");
    }
    b->e_steps.push_back(std::move(step));
  }
  b->e_ntasks = task;
  return GPRM_OK;
}

// The launch steps on b->stream (capturable: kernels only).
int edge_enqueue(gprm_batch* b) {
  const int S = (int)b->ctx.size();
  hipStream_t st = b->stream;
  unsigned long long* clk = b->e_clock.as<unsigned long long>();
  BHIP(b, launch_fill_u32(b->e_ctl.as<unsigned>(), 0u, 3LL * std::max(1, b->e_ntasks), st), "fill");
  BHIP(b, launch_fill_u32(reinterpret_cast<unsigned*>(clk), 0u, 2LL * S, st), "fill");
  BHIP(b, launch_fill_u32(reinterpret_cast<unsigned*>(clk + S), 0xffffffffu, 2, st), "fill");
  for (const auto& step : b->e_steps) {
    const int nt = (int)step->tasks.size();
    BHIP(b, launch_edge_compact_multi(step->d_tasks.as<EdgeTask>(), nt, step->cblocks, st), "compaction launch");
    BHIP(b, launch_edge_fk_multi(step->d_tasks.as<EdgeTask>(), nt, step->fblocks, b->ndof, st),
         "kinematics launch");
    for (int i = 0; i < kBuckets; ++i) {
      if (step->nsva[i] == 0) continue;
      const Bucket& bk = b->buckets[i];
      BHIP(b, launch_sva_multi(step->sva[i].as<SceneTask>(), step->nsva[i], 1, bk.shape, bk.lds, clk + S, nullptr,
                               st),
           "separating-vector launch");
    }
  }
  return GPRM_OK;
}

int edge_run(gprm_batch* b, const EdgeSignature& sg, bool use_graph, float* done_ms, long long* samples) {
  BHIP(b, hipSetDevice(b->device), "hipSetDevice");
  const int S = (int)b->ctx.size();
  if (!(b->have_esig && b->esig == sg)) {
    drop_edge_plan(b);
    b->have_esig = false;
    int rc = edge_prepare(b, sg);
    if (rc != GPRM_OK) return rc;
    b->esig = sg;
    b->have_esig = true;
  }
  // prologue (edge results zeroed, testnum, order, largest testnum per scene)
  hipStream_t st = b->stream;
  BHIP(b, launch_fill_u32(b->e_max.as<unsigned>(), 0u, S, st), "fill");
  BHIP(b, launch_fill_u32(b->e_total.as<unsigned>(), 0u, 2LL * S, st), "fill");
  BHIP(b, launch_edge_prologue_multi(b->e_pro.as<EdgeTask>(), b->e_pro_tasks, b->e_pro_blocks, st),
       "edge prologue launch");
  std::vector<int> maxT(S);
  std::vector<unsigned long long> tot(S);
  BHIP(b, hipMemcpyAsync(maxT.data(), b->e_max.p, sizeof(int) * S, hipMemcpyDeviceToHost, st), "copy");
  BHIP(b, hipMemcpyAsync(tot.data(), b->e_total.p, sizeof(unsigned long long) * S, hipMemcpyDeviceToHost, st),
       "copy");
  BHIP(b, hipStreamSynchronize(st), "sync");
  if (samples) {
    *samples = 0;
    for (int s = 0; s < S; ++s) *samples += (long long)tot[s];
  }
  if (!(b->have_eplan && b->e_plan_key == maxT)) {
    drop_edge_plan(b);
    int rc = edge_plan(b, sg, maxT);
    if (rc != GPRM_OK) return rc;
    b->e_plan_key = maxT;
    b->have_eplan = true;
  }
  if (use_graph) {
    if (!b->eexec) {
      BHIP(b, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
      int rc = edge_enqueue(b);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc != GPRM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return bhip(b, e, "end capture");
      b->egraph = g;
      BHIP(b, hipGraphInstantiate(&b->eexec, b->egraph, nullptr, nullptr, 0), "instantiate graph");
    }
    BHIP(b, hipGraphLaunch(b->eexec, st), "graph launch");
  } else {
    int rc = edge_enqueue(b);
    if (rc != GPRM_OK) return rc;
  }
  BHIP(b, hipStreamSynchronize(st), "sync");
  if (done_ms) {
    std::vector<unsigned long long> td(S + 1);
    BHIP(b, hipMemcpy(td.data(), b->e_clock.p, sizeof(unsigned long long) * (S + 1), hipMemcpyDeviceToHost),
         "copy clocks");
    const unsigned long long t0 = td[S];
    for (int s = 0; s < S; ++s)
      done_ms[s] = (sg.n[s] == 0 || td[s] == 0 || td[s] < t0) ? 0.0f
                                                               : (float)((double)(td[s] - t0) / b->clock_khz);
  }
  return GPRM_OK;
}

}  // namespace

extern "C" {

int gprm_batch_create(gprm_ctx* const* ctxs, int nscenes, gprm_batch** out) {
  if (!ctxs || nscenes < 1 || !out) return bfail(nullptr, GPRM_E_INVALID, "bad arguments");
  *out = nullptr;
  for (int s = 0; s < nscenes; ++s)
    if (!ctxs[s] || !ctxs[s]->inited) return bfail(nullptr, GPRM_E_STATE, "scene context not initialised");
  const int dev = ctxs[0]->device, nd = ctxs[0]->scene.ndof;
  for (int s = 1; s < nscenes; ++s)
    if (ctxs[s]->device != dev || ctxs[s]->scene.ndof != nd)
      return bfail(nullptr, GPRM_E_INVALID, "scenes of a batch need the same device and ndof");
  auto* b = new gprm_batch();
  b->ctx.assign(ctxs, ctxs + nscenes);
  b->device = dev;
  b->ndof = nd;
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->clock_khz, hipDeviceAttributeWallClockRate, dev);
  for (int i = 0; i < kBuckets && e == hipSuccess; ++i) {
    Bucket& bk = b->buckets[i];
    bk.lo = kBucketEdges[i];
    bk.hi = kBucketEdges[i + 1];
    e = hipEventCreate(&bk.ev0);
    if (e == hipSuccess) e = hipEventCreate(&bk.ev1);
  }
  if (e != hipSuccess || b->clock_khz <= 0) {
    int rc = bhip(nullptr, e != hipSuccess ? e : hipErrorInvalidValue, "batch set-up");
    delete b;
    return rc;
  }
  for (int s = 0; s < nscenes; ++s) {
    const int nv = ctxs[s]->img.sumn;
    for (auto& bk : b->buckets)
      if (nv >= bk.lo && nv < bk.hi) {
        bk.scenes.push_back(s);
        bk.lds = std::max(bk.lds, (size_t)ctxs[s]->img.bytes);
      }
  }
  for (auto& bk : b->buckets) {
    if (bk.scenes.empty()) continue;
    // default: five or four scene copies per CU when LDS holds them, else one
    const size_t cap = (size_t)sva_max_lds_bytes() - 16;
    if (5 * (bk.lds + sva_ring_bytes(256)) <= cap) { bk.shape.block = 256; bk.shape.waves = 5; }
    else if (4 * (bk.lds + sva_ring_bytes(256)) <= cap) { bk.shape.block = 256; bk.shape.waves = 4; }
    else { bk.shape.block = 1024; bk.shape.waves = 4; }
    bk.shape.ppu = 64;
    if (bk.lds + sva_ring_bytes(1024) > cap) {
      // every scene of a batch is staged in LDS (single-scene contexts read
      // larger images from HBM)
      g_batch_error = "scene image of " + std::to_string(bk.lds) + " bytes beyond LDS for the multi-scene kernel";
      delete b;
      return GPRM_E_CAPACITY;
    }
    e = sva_multi_shape_finish(bk.lds, &bk.shape, dev);
    if (e != hipSuccess) {
      int rc = bhip(nullptr, e, "scene too large for the multi-scene kernel");
      delete b;
      return rc;
    }
  }
  *out = b;
  return GPRM_OK;
}

void gprm_batch_destroy(gprm_batch* b) {
  if (b) (void)hipSetDevice(b->device);
  delete b;
}

int gprm_batch_bucket_info(gprm_batch* b, int i, gprm_bucket_info* out) {
  if (!b || !out || i < 0 || i >= kBuckets) return bfail(b, GPRM_E_INVALID, "bad arguments");
  const Bucket& bk = b->buckets[i];
  out->min_vertices = bk.lo;
  out->max_vertices = bk.hi;
  out->scenes = (int)bk.scenes.size();
  out->lds_bytes = (int)bk.lds;
  out->shape = gprm_launch_shape{bk.shape.block, bk.shape.waves, bk.shape.ppu, bk.shape.grid};
  return GPRM_OK;
}

int gprm_batch_set_bucket_shape(gprm_batch* b, int i, const gprm_launch_shape* sh) {
  if (!b || !sh || i < 0 || i >= kBuckets) return bfail(b, GPRM_E_INVALID, "bad arguments");
  Bucket& bk = b->buckets[i];
  if (bk.scenes.empty()) return GPRM_OK;
  SvaShape s;
  s.block = sh->block;
  s.waves = sh->waves;
  s.ppu = sh->pairs_per_unit;
  BHIP(b, hipSetDevice(b->device), "hipSetDevice");
  hipError_t e = sva_multi_shape_finish(bk.lds, &s, b->device);
  if (e == hipErrorInvalidValue || e == hipErrorInvalidConfiguration)
    return bfail(b, GPRM_E_INVALID, "launch shape not available for this bucket");
  if (e != hipSuccess) return bhip(b, e, "launch shape");
  bk.shape = s;
  drop_graph(b);  // task tables and the graph depend on the shape
  return GPRM_OK;
}

int gprm_batch_indicator_dev(gprm_batch* b, const float* const* d_q, const long long* offset, const int* n,
                             int* const* d_res, int use_graph, float* done_ms) {
  if (!b) return bfail(nullptr, GPRM_E_INVALID, "null batch");
  Signature sg;
  int rc = signature(b, d_q, offset, n, d_res, &sg);
  if (rc != GPRM_OK) return rc;
  return run(b, sg, use_graph != 0, false, done_ms);
}

int gprm_batch_autotune(gprm_batch* b, const float* const* d_q, const long long* offset, const int* n,
                        int* const* d_res, int reps) {
  if (!b) return bfail(nullptr, GPRM_E_INVALID, "null batch");
  if (reps < 1) return bfail(b, GPRM_E_INVALID, "reps < 1");
  Signature sg;
  int rc = signature(b, d_q, offset, n, d_res, &sg);
  if (rc != GPRM_OK) return rc;
  const gprm_launch_shape variants[] = {{256, 5, 0, 0}, {256, 4, 0, 0}, {1024, 4, 0, 0}};
  const int ppus[] = {32, 64, 128, 256};
  for (int i = 0; i < kBuckets; ++i) {
    Bucket& bk = b->buckets[i];
    if (bk.scenes.empty()) continue;
    gprm_launch_shape best{bk.shape.block, bk.shape.waves, bk.shape.ppu, bk.shape.grid};
    float best_ms = INFINITY;
    for (const auto& v : variants)
      for (int ppu : ppus) {
        gprm_launch_shape sh = v;
        sh.pairs_per_unit = ppu;
        if (gprm_batch_set_bucket_shape(b, i, &sh) != GPRM_OK) continue;  // does not fit this bucket
        // the median of reps timed runs after one warm-up: a single run's
        // noise picked different shapes from run to run
        std::vector<float> ts;
        for (int r = 0; r <= reps; ++r) {
          rc = run(b, sg, false, true, nullptr);
          if (rc != GPRM_OK) return rc;
          float t = 0.0f;
          BHIP(b, hipEventElapsedTime(&t, bk.ev0, bk.ev1), "event time");
          if (r > 0) ts.push_back(t);
        }
        std::sort(ts.begin(), ts.end());
        const float ms = ts[ts.size() / 2];
        if (ms < best_ms) {
          best_ms = ms;
          best = gprm_launch_shape{bk.shape.block, bk.shape.waves, bk.shape.ppu, bk.shape.grid};
        }
      }
    rc = gprm_batch_set_bucket_shape(b, i, &best);
    if (rc != GPRM_OK) return rc;
  }
  return GPRM_OK;
}

int gprm_batch_indicator2_dev(gprm_batch* b, const float* const* d_qs, const float* const* d_qe,
                              const long long* offset, const int* n, int* const* d_res, int use_graph, float* done_ms,
                              long long* samples_total) {
  if (!b) return bfail(nullptr, GPRM_E_INVALID, "null batch");
  EdgeSignature sg;
  int rc = edge_signature(b, d_qs, d_qe, offset, n, d_res, &sg);
  if (rc != GPRM_OK) return rc;
  return edge_run(b, sg, use_graph != 0, done_ms, samples_total);
}

const char* gprm_batch_last_error(gprm_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

}  // extern "C"
