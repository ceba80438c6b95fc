// prm.cpp -- single-process bidirectional PRM on one GPU.
//
// The planner of the reference (PRMSolver<ndof>, prm/src/lib/prm/prmsolver.
// {h,cpp}) restated for one process and one device.  Its process 5
// (prmsolver.cpp:1765-1895) runs two workers: while one worker's edge checks
// are in flight, the host inserts the other worker's results and prepares
// its next step.  Here a worker is a HIP stream with its own buffers, and a
// step is device work from end to end (planner.hip):
//   part 1  weights, sampling around weighted base nodes with the exact
//           point check, selection of the new nodes, radius search grouped
//           by (new node, tree) (processing_step_part1a, :1973-2040);
//           its counts are read back (the only read-back of a step) to size
//   part 2  the edge checks: [edge budget ranking], endpoint gather, exact
//           edge checks (indicator2_async, :2015; wait, :2071);
//   part 3  insertion: every new node with a free edge joins that tree with
//           all its free edges; the first node joining both trees connects
//           them (calc_conn + part3a, :2076-2187); cell indices rebuilt.
// The trees live on the device.  Parts 1 and 3, which read and write them,
// run in the order the host issues them (one event chains them across the
// workers' streams), so a worker sees every node inserted before its part 1
// was issued, as in the reference's single-threaded host loop; part 2 of one
// worker runs beside the other worker's parts 1 and 3.  The host reads the
// trees back once, at the end, and runs Dijkstra on each tree
// (do_dijkstra, :2804-2846).
// The optional model (surrogate.hip) ranks; it never decides.  Sampling bias
// drops candidates before their exact check, edge ranking chooses which
// edges are checked; every node and edge that enters a tree has been decided
// free by the exact checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "../../include/gprm.h"
#include "context.h"
#include "planner.h"
#include "runtime.h"
#include "scene.h"
#include "surrogate.h"

namespace gprm {
namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

// Points scored along an edge by the model: interior points at t = s/(S+1).
// One point, the midpoint: on the 6-DOF scene three points (t = 1/4, 1/2,
// 3/4) cost the planner 1.0 ms of inference per run at 4,096 nodes per step
// and bought no fewer steps or edge samples.
constexpr int kScorePoints = 1;
// Undirected edges a tree may store, per node of its capacity.
constexpr long long kEdgesPerNode = 16;

struct DevArr {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevArr() { if (p) (void)hipFree(p); }
  bool ensure(size_t b) {
    if (b <= bytes) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, b) != hipSuccess) return false;
    bytes = b;
    return true;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// One tree on the host after the search (the reference's `graph`,
// prmsolver.h:40-58, without the block index).
struct Tree {
  int ndof = 0, n = 0;
  std::vector<float> q;                 // AoS ndof*n
  std::vector<int> surrnum;             // neighbours per node
  // adjacency as directed entries (both directions of an edge); csr()
  // groups them by node, keeping each node's order
  std::vector<int> ef, et;
  std::vector<float> ew;
};

// Adjacency of a tree in CSR form: node l's entries (neighbour, weight) at
// [off[l], off[l+1]), in the order they were added.
struct Csr {
  std::vector<int> off, to;
  std::vector<float> w;
};
Csr csr(const Tree& t) {
  Csr c;
  c.off.assign((size_t)t.n + 1, 0);
  for (int f : t.ef) ++c.off[(size_t)f + 1];
  for (size_t l = 1; l < c.off.size(); ++l) c.off[l] += c.off[l - 1];
  std::vector<int> cur(c.off.begin(), c.off.end() - 1);
  c.to.resize(t.ef.size());
  c.w.resize(t.ef.size());
  for (size_t k = 0; k < t.ef.size(); ++k) {
    const int x = cur[t.ef[k]]++;
    c.to[x] = t.et[k];
    c.w[x] = t.ew[k];
  }
  return c;
}

// Dijkstra on one tree (do_dijkstra, prmsolver.cpp:2804-2846).
bool dijkstra(const Tree& t, const Csr& g, int from, int to, std::vector<int>* path) {
  std::vector<float> dist(t.n, INFINITY);
  std::vector<int> parent(t.n, -1);
  typedef std::pair<float, int> QE;
  std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
  dist[from] = 0.0f;
  parent[from] = from;
  pq.push({0.0f, from});
  while (!pq.empty()) {
    QE e = pq.top();
    pq.pop();
    if (e.first > dist[e.second]) continue;
    if (e.second == to) {
      path->assign(1, to);
      for (int p = to; p != from;) {
        p = parent[p];
        path->push_back(p);
      }
      return true;
    }
    for (int x = g.off[e.second]; x < g.off[e.second + 1]; ++x) {
      const int nb = g.to[x];
      float nd = e.first + g.w[x];
      if (nd < dist[nb]) {
        dist[nb] = nd;
        parent[nb] = e.second;
        pq.push({nd, nb});
      }
    }
  }
  return false;
}

template <class T>
bool write_bin(const std::string& path, const T* v, size_t n) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) return false;
  f.write(reinterpret_cast<const char*>(v), (std::streamsize)(sizeof(T) * n));
  return (bool)f;
}

// One tree in the reference's output layout (store_graph,
// prmsolver.cpp:2976-3055, read by prm/matlab/lib/prmoutputread.m).  The
// reference's trees are block-indexed with gaps; these are dense, so the
// compressed files (...c.bin) equal the raw ones and the index mapping is the
// identity.  Edge weights are written as int, truncated, as the reference's
// std::vector<int> edgesweight stores them.  blocknum.bin ends up holding the
// block count (the reference writes it twice; the second write wins): one
// block holding every node.
bool store_tree(const std::string& dir, const Tree& t, const Csr& g, const std::vector<int>* path, int start,
                int end) {
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) return false;
  std::vector<int> from, to, weight;
  for (int l = 0; l < t.n; ++l)
    for (int x = g.off[l]; x < g.off[l + 1]; ++x) {
      from.push_back(l);
      to.push_back(g.to[x]);
      weight.push_back((int)g.w[x]);
    }
  const int n = t.n, zero = 0, one = 1;
  bool ok = write_bin(dir + "/qstorage.bin", t.q.data(), (size_t)t.ndof * n) &&
            write_bin(dir + "/surrnum.bin", t.surrnum.data(), (size_t)n) &&
            write_bin(dir + "/edgesfrom.bin", from.data(), from.size()) &&
            write_bin(dir + "/edgesto.bin", to.data(), to.size()) &&
            write_bin(dir + "/edgesfromc.bin", from.data(), from.size()) &&
            write_bin(dir + "/edgestoc.bin", to.data(), to.size()) &&
            write_bin(dir + "/edgesweight.bin", weight.data(), weight.size()) &&
            write_bin(dir + "/qstoragec.bin", t.q.data(), (size_t)t.ndof * n) &&
            write_bin(dir + "/blockpos.bin", &zero, 1) && write_bin(dir + "/blocknum.bin", &one, 1) &&
            write_bin(dir + "/newblockpos.bin", &n, 1) && write_bin(dir + "/numc.bin", &n, 1) &&
            write_bin(dir + "/startc.bin", &start, 1) && write_bin(dir + "/endc.bin", &end, 1);
  if (ok && path)
    ok = write_bin(dir + "/path.bin", path->data(), path->size()) &&
         write_bin(dir + "/pathc.bin", path->data(), path->size());
  return ok;
}

// Device state of one tree (PlanTree's arrays).
struct DevTree {
  DevArr q, surr, cum, bsum, efrom, eto, ew, cell_id, counts, starts, cursor, cbsum, sorted, sid;
};

// What part 1 reads back: the step's counts and the planner's totals.
struct Readback {
  StepInfo info;
  PlanState state;
};

// HIP events of one timed interval on a worker stream, folded into a total
// once the solve has finished.
struct Interval {
  hipEvent_t a, b;
  double* total;
};

// One worker of the process-5 scheme: a stream (its slot's), the buffers of
// one step, and the step's state between its parts.
struct Worker {
  int slot = -1;
  Slot* sl = nullptr;
  hipStream_t st = nullptr;
  DevArr cand, codes, pskip, pprob, qnew, info, tile, cnt, off, sc_off, enode, ed2, egroup, eskip, escore, er, qs,
      qe, free_cnt, pos, spts, sprob;
  Readback* pin = nullptr;
  hipEvent_t ev_rb = nullptr;
  int E = 0, S = 0, batch = 0;
  ~Worker() {
    if (ev_rb) (void)hipEventDestroy(ev_rb);
    if (pin) (void)hipHostFree(pin);
  }
};

}  // namespace

// Device state of the planner kept by a context between solves, so that a
// solve allocates only when it needs more than an earlier one did (the
// 6-DOF surrogate comparison's solves take about 10 ms).
struct PlanCache {
  DevArr state;
  DevTree dt[2];
  std::vector<std::unique_ptr<Worker>> workers;
  DevArr dist, parent, flags, chain, lens, pcoords;  // path extraction
  std::vector<hipEvent_t> timing;                    // interval events, reused
  hipEvent_t ev_chain = nullptr;
  ~PlanCache() {
    (void)hipDeviceSynchronize();
    for (hipEvent_t e : timing) (void)hipEventDestroy(e);
    if (ev_chain) (void)hipEventDestroy(ev_chain);
  }
};

}  // namespace gprm

using namespace gprm;

extern "C" int gprm_prm_solve(gprm_ctx* ctx, const float* qstart, const float* qgoal, const gprm_prm_params* prm,
                              gprm_prm_result* res, float* path, int path_cap) {
  if (!ctx || !qstart || !qgoal || !prm || !res) return GPRM_E_INVALID;
  std::memset(res, 0, sizeof(*res));
  const int nd = gprm_dim(ctx);
  if (prm->num < 2 || prm->num > (1 << 14) || !(prm->D > 0.0f) || prm->max_nodes < 2 || prm->max_steps < 1 ||
      prm->nbuf < 1 || prm->edge_budget < 0 || prm->workers < 0 || prm->workers > 2)
    return GPRM_E_INVALID;
  const auto t0 = Clock::now();
  // start and goal must be free (prm/src/main.cpp:135-138)
  int c0 = 0, c1 = 0;
  int rc = gprm_indicator1(ctx, qstart, &c0);
  if (rc == GPRM_OK) rc = gprm_indicator1(ctx, qgoal, &c1);
  if (rc != GPRM_OK) return rc;
  res->point_checks += 2;
  if (c0 != 0 || c1 != 0) return GPRM_OK;  // found = 0

  std::unique_lock<std::mutex> lock(ctx->mu);
  // the planner keeps its slots busy across calls; the profiling events of
  // the check calls are not harvested inside a solve
  struct ProfileOff {
    gprm_ctx* c;
    bool was;
    ~ProfileOff() { c->profile = was; }
  } prof_off{ctx, ctx->profile};
  ctx->profile = false;
  if (hipSetDevice(ctx->device) != hipSuccess) return GPRM_E_HIP;
  Surrogate* model = static_cast<Surrogate*>(prm->surrogate);
  const bool bias = model && prm->bias_sampling;
  const bool rank_by_model = model && prm->order_edges;
  const int budget = prm->edge_budget;
  const int W = prm->workers == 0 ? 1 : prm->workers;
  const int num = prm->num, cap = prm->max_nodes, nbuf = prm->nbuf;
  const int groups = 2 * num;
  const float D = prm->D, D2 = D * D;
  // every candidate edge is shorter than D (the radius search's own float
  // sum, which the edge prologue repeats), so testnum <= D/dq + 1
  const int maxT_bound = (int)(D / ctx->dq * (1.0 + 1e-3)) + 2;
  const int batch_max = 32 * num;
  std::vector<float> lo(nd), hi(nd);
  for (int i = 0; i < nd; ++i) {
    lo[i] = ctx->scene.mins[i];
    hi[i] = ctx->scene.maxs[i];
  }
  const CellGrid grid = make_cell_grid(nd, lo.data(), hi.data(), D);
  const int nb = plan_scan_blocks(cap), nbc = plan_scan_blocks(grid.ncells);
  const long long ecap_ll = std::min<long long>(kEdgesPerNode * cap, (1LL << 31) - 1);
  const int ecap = (int)ecap_ll;

  // ---- device state (kept by the context between solves) ------------------------
  if (!ctx->plan_cache) ctx->plan_cache = std::make_shared<PlanCache>();
  PlanCache& pc = *ctx->plan_cache;
  DevArr& d_state = pc.state;
  DevTree* dt = pc.dt;
  PlanTree T[2];
  if (!d_state.ensure(sizeof(PlanState))) return GPRM_E_HIP;
  PlanState* state = d_state.as<PlanState>();
  for (int g = 0; g < 2; ++g) {
    DevTree& d = dt[g];
    if (!d.q.ensure(sizeof(float) * nd * (size_t)cap) || !d.surr.ensure(sizeof(int) * (size_t)cap) ||
        !d.cum.ensure(sizeof(double) * (size_t)cap) || !d.bsum.ensure(sizeof(double) * (size_t)(nb + 1)) ||
        !d.efrom.ensure(sizeof(int) * (size_t)ecap) || !d.eto.ensure(sizeof(int) * (size_t)ecap) ||
        !d.ew.ensure(sizeof(float) * (size_t)ecap) || !d.cell_id.ensure(sizeof(int) * (size_t)cap) ||
        !d.counts.ensure(sizeof(int) * (size_t)grid.ncells) || !d.starts.ensure(sizeof(int) * (size_t)grid.ncells) ||
        !d.cursor.ensure(sizeof(int) * (size_t)grid.ncells) || !d.cbsum.ensure(sizeof(int) * (size_t)(nbc + 1)) ||
        !d.sorted.ensure(sizeof(float) * nd * (size_t)cap) || !d.sid.ensure(sizeof(int) * (size_t)cap))
      return GPRM_E_HIP;
    PlanTree& t = T[g];
    t.q = d.q.as<float>();
    t.surr = d.surr.as<int>();
    t.n = &state->n[g];
    t.cum = d.cum.as<double>();
    t.bsum = d.bsum.as<double>();
    t.efrom = d.efrom.as<int>();
    t.eto = d.eto.as<int>();
    t.ew = d.ew.as<float>();
    t.ecount = &state->ecount[g];
    t.ecap = ecap;
    t.cell_id = d.cell_id.as<int>();
    t.counts = d.counts.as<int>();
    t.starts = d.starts.as<int>();
    t.cursor = d.cursor.as<int>();
    t.cbsum = d.cbsum.as<int>();
    t.sorted = d.sorted.as<float>();
    t.sid = d.sid.as<int>();
  }

  // ---- workers ------------------------------------------------------------------
  std::vector<std::unique_ptr<Worker>>& wk = pc.workers;
  while ((int)wk.size() < W) wk.emplace_back(new Worker());
  // the slots are held for the solve and released, with their streams
  // drained, however it ends
  struct Slots {
    gprm_ctx* c;
    std::vector<std::unique_ptr<Worker>>* w;
    int W;
    ~Slots() {
      for (int i = 0; i < W; ++i) {
        Worker& x = *(*w)[i];
        if (x.st) (void)hipStreamSynchronize(x.st);
        if (x.slot >= 0) slot_release(c, x.slot);
        x.slot = -1;
        x.sl = nullptr;
        x.st = nullptr;
      }
    }
  } slots_guard{ctx, &wk, W};
  for (int w = 0; w < W; ++w) {
    Worker& x = *wk[w];
    x.E = x.S = x.batch = 0;
    if ((rc = slot_acquire(ctx, &x.slot)) != GPRM_OK) return rc;
    x.sl = ctx->slots[x.slot].get();
    x.st = x.sl->stream;
    if ((!x.pin && hipHostMalloc((void**)&x.pin, sizeof(Readback), hipHostMallocDefault) != hipSuccess) ||
        (!x.ev_rb && hipEventCreateWithFlags(&x.ev_rb, hipEventDisableTiming) != hipSuccess))
      return GPRM_E_HIP;
    if (!x.cand.ensure(sizeof(float) * nd * (size_t)batch_max) || !x.codes.ensure(sizeof(int) * (size_t)batch_max) ||
        !x.qnew.ensure(sizeof(float) * nd * (size_t)num) || !x.info.ensure(sizeof(StepInfo)) ||
        !x.cnt.ensure(sizeof(int) * (size_t)groups) || !x.off.ensure(sizeof(int) * (size_t)(groups + 1)) ||
        !x.sc_off.ensure(sizeof(int) * (size_t)(groups + 1)) || !x.enode.ensure(sizeof(int) * (size_t)nbuf) ||
        !x.ed2.ensure(sizeof(float) * (size_t)nbuf) || !x.egroup.ensure(sizeof(int) * (size_t)nbuf) ||
        !x.er.ensure(sizeof(int) * (size_t)nbuf) || !x.qs.ensure(sizeof(float) * nd * (size_t)nbuf) ||
        !x.qe.ensure(sizeof(float) * nd * (size_t)nbuf) || !x.free_cnt.ensure(sizeof(int) * (size_t)groups) ||
        !x.pos.ensure(sizeof(int) * (size_t)groups) ||
        !x.tile.ensure(sizeof(int) * 3 * (size_t)(batch_max / 1024 + 1)) ||
        (bias && (!x.pskip.ensure((size_t)batch_max) || !x.pprob.ensure(sizeof(float) * (size_t)batch_max))) ||
        (budget > 0 && (!x.eskip.ensure((size_t)nbuf) || !x.escore.ensure(sizeof(float) * (size_t)nbuf))))
      return GPRM_E_HIP;
    if ((rc = reserve_slot(ctx, *x.sl, batch_max, 0, maxT_bound)) != GPRM_OK) return rc;
  }

  // ---- initial trees: the start and the goal, indexed ---------------------------------
  {
    hipStream_t st = wk[0]->st;
    PlanState init{};
    init.n[0] = init.n[1] = 1;
    if (hipMemcpyAsync(state, &init, sizeof(init), hipMemcpyHostToDevice, st) != hipSuccess) return GPRM_E_HIP;
    const float* roots[2] = {qstart, qgoal};
    for (int g = 0; g < 2; ++g) {
      for (int i = 0; i < nd; ++i)
        if (hipMemcpyAsync(T[g].q + (size_t)cap * i, roots[g] + i, sizeof(float), hipMemcpyHostToDevice, st) !=
            hipSuccess)
          return GPRM_E_HIP;
      if (launch_fill_i32(T[g].surr, 0, 1, st) != hipSuccess) return GPRM_E_HIP;
    }
    if (launch_plan_index(nd, T[0], T[1], cap, grid, st) != hipSuccess) return GPRM_E_HIP;
    // the host copies above read host memory when the stream runs them
    if (hipStreamSynchronize(st) != hipSuccess) return GPRM_E_HIP;
  }
  if (!pc.ev_chain && hipEventCreateWithFlags(&pc.ev_chain, hipEventDisableTiming) != hipSuccess) return GPRM_E_HIP;
  hipEvent_t ev_chain = pc.ev_chain;
  if (hipEventRecord(ev_chain, wk[0]->st) != hipSuccess) return GPRM_E_HIP;

  // timing: pairs of events per interval (from the context's pool), folded
  // at the end
  std::vector<Interval> timed;
  // begin_iv returns the interval's index (-1 on failure), end_iv closes it
  auto begin_iv = [&](hipStream_t st, double* total) -> int {
    const size_t k = 2 * timed.size();
    while (pc.timing.size() < k + 2) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      pc.timing.push_back(e);
    }
    timed.push_back(Interval{pc.timing[k], pc.timing[k + 1], total});
    return hipEventRecord(timed.back().a, st) == hipSuccess ? (int)timed.size() - 1 : -1;
  };
  auto end_iv = [&](int i, hipStream_t st) -> bool { return i >= 0 && hipEventRecord(timed[i].b, st) == hipSuccess; };

  Rng rng(prm->seed * 0x9E3779B97F4A7C15ULL + 12345);
  const int want0 = num / 2, want1 = num - num / 2;
  int batch = 4 * num;
  double ms_wait = 0.0;
  long long predictions = 0;
  bool stop = false;

  // part 1: sampling, selection, radius search; reads back the counts
  auto part1 = [&](Worker& w) -> int {
    hipStream_t st = w.st;
    if (hipStreamWaitEvent(st, ev_chain, 0) != hipSuccess) return GPRM_E_HIP;
    w.batch = batch;
    const unsigned long long key = rng.next();
    StepInfo* info = w.info.as<StepInfo>();
    const int iv_s = begin_iv(st, &res->ms_sampling);
    if (iv_s < 0) return GPRM_E_HIP;
    if (launch_plan_weights(T[0], T[1], nb, st) != hipSuccess ||
        launch_plan_sample(nd, w.batch, T[0], T[1], cap, nb, D, key, w.cand.as<float>(), st) != hipSuccess)
      return GPRM_E_HIP;
    const unsigned char* pskip = nullptr;
    if (bias) {
      // every second pair of candidates must be predicted free; the others,
      // and all candidates without a model, go to the exact check
      const int iv_m = begin_iv(st, &res->ms_model);
      if (iv_m < 0 || surrogate_predict(model, w.cand.as<float>(), w.batch, w.batch, w.pprob.as<float>(), st) != 0 ||
          !end_iv(iv_m, st) ||
          launch_plan_bias_mask(w.pprob.as<float>(), w.batch, w.pskip.as<unsigned char>(), st) != hipSuccess)
        return GPRM_E_HIP;
      predictions += w.batch;
      pskip = w.pskip.as<unsigned char>();
    }
    int r = enqueue_points(ctx, *w.sl, w.cand.as<float>(), w.batch, w.batch, w.codes.as<int>(), st, pskip);
    if (r != GPRM_OK) return r;
    if (launch_plan_select(nd, w.batch, w.codes.as<int>(), pskip, w.cand.as<float>(), want0, want1, num,
                           w.qnew.as<float>(), w.tile.as<int>(), info, state, st) != hipSuccess ||
        !end_iv(iv_s, st))
      return GPRM_E_HIP;
    // with an edge budget the radius search stores each group's edges
    // contiguously (the budget ranks within groups)
    const int iv_n = begin_iv(st, &res->ms_near);
    if (iv_n < 0 ||
        launch_plan_near(nd, num, w.qnew.as<float>(), info, T[0], T[1], cap, grid, D2, w.cnt.as<int>(),
                         w.off.as<int>(), w.sc_off.as<int>(), nbuf, budget, w.enode.as<int>(),
                         w.ed2.as<float>(), w.egroup.as<int>(), st) != hipSuccess ||
        !end_iv(iv_n, st) || hipEventRecord(ev_chain, st) != hipSuccess ||
        hipMemcpyAsync(&w.pin->info, info, sizeof(StepInfo), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&w.pin->state, state, sizeof(PlanState), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(w.ev_rb, st) != hipSuccess)
      return GPRM_E_HIP;
    const auto tw = Clock::now();
    if (hipEventSynchronize(w.ev_rb) != hipSuccess) return GPRM_E_HIP;
    ms_wait += ms_since(tw);
    const StepInfo& si = w.pin->info;
    w.E = si.E;
    w.S = si.S;
    // the next batch: enough candidates for 1.5 num free ones at the rate
    // just seen, at least 2 num
    const double rate = std::max(1e-3, (double)si.accepted / std::max(1, w.batch));
    int want_batch = (int)std::ceil(1.5 * num / rate);
    if (si.nnew < num) want_batch = std::max(want_batch, 2 * w.batch);
    batch = std::min(batch_max, std::max(2 * num, (want_batch + 255) / 256 * 256));
    if (w.pin->state.conn || w.pin->state.full || w.pin->state.over_bound) stop = true;
    return GPRM_OK;
  };

  // part 2: edge checks of the step (beside the other worker's parts)
  auto part2 = [&](Worker& w) -> int {
    hipStream_t st = w.st;
    const int E = w.E;
    if (launch_fill_i32(w.free_cnt.as<int>(), 0, groups, st) != hipSuccess) return GPRM_E_HIP;
    if (E == 0) return GPRM_OK;
    const int iv_e = begin_iv(st, &res->ms_edges);
    if (iv_e < 0) return GPRM_E_HIP;
    const unsigned char* eskip = nullptr;
    if (budget > 0) {
      const float* prob = nullptr;
      if (rank_by_model && w.S > 0) {
        const long long np = (long long)w.S * kScorePoints;
        if (w.spts.bytes < sizeof(float) * nd * (size_t)np || w.sprob.bytes < sizeof(float) * (size_t)np) {
          // work in flight may use the old buffers
          if (hipDeviceSynchronize() != hipSuccess || !w.spts.ensure(sizeof(float) * nd * (size_t)np * 3 / 2) ||
              !w.sprob.ensure(sizeof(float) * (size_t)np * 3 / 2))
            return GPRM_E_HIP;
        }
        const int iv_m = begin_iv(st, &res->ms_model);
        if (iv_m < 0 ||
            launch_plan_score_points(nd, E, nbuf, budget, kScorePoints, w.cnt.as<int>(), w.off.as<int>(),
                                     w.sc_off.as<int>(), w.egroup.as<int>(), w.enode.as<int>(), w.qnew.as<float>(), num,
                                     T[0], T[1], cap, w.spts.as<float>(), np, st) != hipSuccess ||
            surrogate_predict(model, w.spts.as<float>(), (int)np, np, w.sprob.as<float>(), st) != 0 ||
            !end_iv(iv_m, st))
          return GPRM_E_HIP;
        predictions += np;
        prob = w.sprob.as<float>();
      }
      if (launch_plan_rank(E, nbuf, budget, kScorePoints, w.cnt.as<int>(), w.off.as<int>(), w.sc_off.as<int>(),
                           w.egroup.as<int>(), w.ed2.as<float>(), prob, w.escore.as<float>(),
                           w.eskip.as<unsigned char>(), state, st) != hipSuccess)
        return GPRM_E_HIP;
      eskip = w.eskip.as<unsigned char>();
    }
    int r = reserve_slot(ctx, *w.sl, w.batch, E, maxT_bound);
    if (r != GPRM_OK) return r;
    if (launch_plan_gather(nd, E, w.egroup.as<int>(), w.enode.as<int>(), w.qnew.as<float>(), num, T[0], T[1], cap,
                           w.qs.as<float>(), w.qe.as<float>(), st) != hipSuccess)
      return GPRM_E_HIP;
    r = enqueue_edges(ctx, *w.sl, w.qs.as<float>(), w.qe.as<float>(), E, E, w.er.as<int>(), st, maxT_bound, eskip);
    if (r != GPRM_OK) return r;
    if (launch_plan_edges_done(E, w.egroup.as<int>(), w.er.as<int>(), eskip, w.free_cnt.as<int>(),
                               edge_sample_total(*w.sl), edge_max_testnum(*w.sl), maxT_bound, state, st) !=
            hipSuccess ||
        !end_iv(iv_e, st))
      return GPRM_E_HIP;
    return GPRM_OK;
  };

  // part 3: insertion and the cell indices (in issue order with the parts 1)
  auto part3 = [&](Worker& w) -> int {
    hipStream_t st = w.st;
    if (hipStreamWaitEvent(st, ev_chain, 0) != hipSuccess) return GPRM_E_HIP;
    const int iv_i = begin_iv(st, &res->ms_insert);
    if (iv_i < 0 ||
        launch_plan_insert(nd, num, w.E, w.info.as<StepInfo>(), w.free_cnt.as<int>(), w.pos.as<int>(),
                           w.qnew.as<float>(), w.egroup.as<int>(), w.enode.as<int>(), w.ed2.as<float>(),
                           w.er.as<int>(), T[0], T[1], cap, state, st) != hipSuccess ||
        launch_plan_index(nd, T[0], T[1], cap, grid, st) != hipSuccess || !end_iv(iv_i, st) ||
        hipEventRecord(ev_chain, st) != hipSuccess)
      return GPRM_E_HIP;
    ++res->steps;
    return GPRM_OK;
  };

  // the process-5 schedule: part 1 and 2 of every worker, then per worker in
  // turn part 3 (its insertion), part 1 and part 2 of its next step
  for (int w = 0; w < W && !stop; ++w) {
    if ((rc = part1(*wk[w])) != GPRM_OK) return rc;
    if (!stop && (rc = part2(*wk[w])) != GPRM_OK) return rc;
  }
  while (!stop && res->steps < prm->max_steps) {
    for (int w = 0; w < W && !stop && res->steps < prm->max_steps; ++w) {
      if ((rc = part3(*wk[w])) != GPRM_OK) return rc;
      if (res->steps >= prm->max_steps) break;
      if ((rc = part1(*wk[w])) != GPRM_OK) return rc;
      if (!stop && (rc = part2(*wk[w])) != GPRM_OK) return rc;
    }
  }

  // ---- the trees back on the host ---------------------------------------------------
  const auto tf = Clock::now();
  if (hipDeviceSynchronize() != hipSuccess) return GPRM_E_HIP;
  for (auto& iv : timed) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, iv.a, iv.b) != hipSuccess) return GPRM_E_HIP;
    *iv.total += ms;
  }
  PlanState fin{};
  if (hipMemcpy(&fin, state, sizeof(fin), hipMemcpyDeviceToHost) != hipSuccess) return GPRM_E_HIP;
  if (fin.over_bound) {
    // some samples of an edge were never scheduled: its decision is not
    // exact, so the trees cannot be trusted
    ctx->err = "planner: an edge had more samples than the radius allows";
    return GPRM_E_STATE;
  }
  res->point_checks += (long long)fin.point_checks;
  res->candidates_skipped = (long long)fin.dropped;
  res->edges_checked = (long long)fin.edges_checked;
  res->edges_free = (long long)fin.edges_free;
  res->edge_samples = (long long)fin.edge_samples;
  res->edges_skipped = (long long)fin.edges_skipped;
  res->model_predictions = predictions;
  res->nodes_left = fin.n[0];
  res->nodes_right = fin.n[1];
  res->ms_wait = ms_wait;
  const int conn_left = fin.conn ? fin.conn_left : -1, conn_right = fin.conn ? fin.conn_right : -1;
  // chains conn_left -> start root (tree 0) and conn_right -> goal root
  // (tree 1); pl = [conn_left, ..., 0], pr = [0, ..., conn_right] as the
  // reference's dijkstra_result paths
  std::vector<int> pl, pr;
  std::vector<float> coords;  // path configurations, AoS
  bool connected = false;
  if (conn_left >= 0) {
    // shortest paths on the device: relaxation rounds in batches of 8, each
    // round with its own changed flags, until a round changes nothing
    const int max_edges = (int)std::min<unsigned>(std::max(fin.ecount[0], fin.ecount[1]), (unsigned)ecap);
    const int maxlen = std::min(cap, 1 << 16);
    constexpr int kRounds = 8;
    DevArr &dist = pc.dist, &parent = pc.parent, &flags = pc.flags, &chain = pc.chain, &lens = pc.lens,
           &pcoords = pc.pcoords;
    hipStream_t st = wk[0]->st;
    if (!dist.ensure(sizeof(int) * 2 * (size_t)cap) || !parent.ensure(sizeof(int) * 2 * (size_t)cap) ||
        !flags.ensure(sizeof(int) * 2 * kRounds) || !chain.ensure(sizeof(int) * 2 * (size_t)maxlen) ||
        !lens.ensure(sizeof(int) * 2) || !pcoords.ensure(sizeof(float) * 2 * (size_t)maxlen * nd))
      return GPRM_E_HIP;
    if (launch_plan_paths_init(T[0], T[1], cap, dist.as<int>(), parent.as<int>(), st) != hipSuccess)
      return GPRM_E_HIP;
    const long long max_rounds = (long long)std::max(fin.n[0], fin.n[1]) + 1;
    int hf[2 * kRounds];
    for (long long r = 0; r < max_rounds; r += kRounds) {
      if (launch_fill_i32(flags.as<int>(), 0, 2 * kRounds, st) != hipSuccess) return GPRM_E_HIP;
      for (int k = 0; k < kRounds; ++k)
        if (launch_plan_paths_relax(T[0], T[1], cap, max_edges, dist.as<int>(), flags.as<int>() + 2 * k, st) !=
            hipSuccess)
          return GPRM_E_HIP;
      if (hipMemcpyAsync(hf, flags.p, sizeof(hf), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return GPRM_E_HIP;
      if (hf[2 * kRounds - 2] == 0 && hf[2 * kRounds - 1] == 0) break;
    }
    int hl[2] = {-1, -1};
    if (launch_plan_paths_finish(nd, T[0], T[1], cap, max_edges, dist.as<int>(), parent.as<int>(), state, maxlen,
                                 chain.as<int>(), lens.as<int>(), pcoords.as<float>(), st) != hipSuccess ||
        hipMemcpyAsync(hl, lens.p, sizeof(hl), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return GPRM_E_HIP;
    if (hl[0] > 0 && hl[1] > 0) {
      std::vector<int> c0(hl[0]), c1(hl[1]);
      std::vector<float> x0((size_t)hl[0] * nd), x1((size_t)hl[1] * nd);
      if (hipMemcpy(c0.data(), chain.p, sizeof(int) * hl[0], hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(c1.data(), chain.as<int>() + maxlen, sizeof(int) * hl[1], hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(x0.data(), pcoords.p, sizeof(float) * x0.size(), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(x1.data(), pcoords.as<float>() + (size_t)maxlen * nd, sizeof(float) * x1.size(),
                    hipMemcpyDeviceToHost) != hipSuccess)
        return GPRM_E_HIP;
      pl = c0;
      pr.assign(c1.rbegin(), c1.rend());
      // start ... conn (tree 0), then conn's successors ... goal (tree 1);
      // conn_left and conn_right hold the same configuration
      for (int i = hl[0] - 1; i >= 0; --i) coords.insert(coords.end(), &x0[(size_t)nd * i], &x0[(size_t)nd * (i + 1)]);
      for (int i = 1; i < hl[1]; ++i) coords.insert(coords.end(), &x1[(size_t)nd * i], &x1[(size_t)nd * (i + 1)]);
      connected = true;
    }
  }
  // the trees on the host: for the output files, or when the device chains
  // broke (a parent lost to float absorption), for Dijkstra on the host
  Tree H[2];
  Csr G[2];
  const bool need_host = prm->output_dir != nullptr || (conn_left >= 0 && !connected);
  if (need_host) {
    for (int g = 0; g < 2; ++g) {
      Tree& t = H[g];
      t.ndof = nd;
      t.n = fin.n[g];
      std::vector<float> soa((size_t)nd * t.n);
      for (int i = 0; i < nd; ++i)
        if (hipMemcpy(soa.data() + (size_t)t.n * i, T[g].q + (size_t)cap * i, sizeof(float) * t.n,
                      hipMemcpyDeviceToHost) != hipSuccess)
          return GPRM_E_HIP;
      t.q.resize((size_t)nd * t.n);
      for (int m = 0; m < t.n; ++m)
        for (int i = 0; i < nd; ++i) t.q[(size_t)nd * m + i] = soa[(size_t)t.n * i + m];
      t.surrnum.resize(t.n);
      const int ne = (int)std::min<unsigned>(fin.ecount[g], (unsigned)ecap);
      std::vector<int> ef(ne), et(ne);
      std::vector<float> ew(ne);
      if (hipMemcpy(t.surrnum.data(), T[g].surr, sizeof(int) * t.n, hipMemcpyDeviceToHost) != hipSuccess ||
          (ne > 0 && (hipMemcpy(ef.data(), T[g].efrom, sizeof(int) * ne, hipMemcpyDeviceToHost) != hipSuccess ||
                      hipMemcpy(et.data(), T[g].eto, sizeof(int) * ne, hipMemcpyDeviceToHost) != hipSuccess ||
                      hipMemcpy(ew.data(), T[g].ew, sizeof(float) * ne, hipMemcpyDeviceToHost) != hipSuccess)))
        return GPRM_E_HIP;
      t.ef.resize(2 * (size_t)ne);
      t.et.resize(2 * (size_t)ne);
      t.ew.resize(2 * (size_t)ne);
      for (int k = 0; k < ne; ++k) {
        t.ef[2 * k] = ef[k];
        t.et[2 * k] = et[k];
        t.ef[2 * k + 1] = et[k];
        t.et[2 * k + 1] = ef[k];
        t.ew[2 * k] = t.ew[2 * k + 1] = ew[k];
      }
      G[g] = csr(t);
    }
    if (conn_left >= 0 && !connected) {
      connected = dijkstra(H[0], G[0], 0, conn_left, &pl) && dijkstra(H[1], G[1], conn_right, 0, &pr);
      if (connected) {
        // pl = [conn_left, ..., start], pr = [goal, ..., conn_right]
        for (int i = (int)pl.size() - 1; i >= 0; --i)
          coords.insert(coords.end(), &H[0].q[(size_t)nd * pl[i]], &H[0].q[(size_t)nd * (pl[i] + 1)]);
        for (int i = (int)pr.size() - 2; i >= 0; --i)
          coords.insert(coords.end(), &H[1].q[(size_t)nd * pr[i]], &H[1].q[(size_t)nd * (pr[i] + 1)]);
      }
    }
  }
  if (connected) {
    res->found = 1;
    res->path_len = (int)(coords.size() / nd);
    if (path) std::memcpy(path, coords.data(), sizeof(float) * nd * (size_t)std::min(path_cap, res->path_len));
  }
  res->ms_final = ms_since(tf);
  res->ms_total = ms_since(t0);
  if (prm->output_dir) {
    // store_results (prmsolver.cpp:2947-2973): graphl from the start root to
    // the connection, graphr from the connection to the goal root
    const std::string root(prm->output_dir);
    const int conn = connected ? 1 : 0;
    std::error_code ec;
    std::filesystem::create_directories(root, ec);
    if (ec || !store_tree(root + "/graphl", H[0], G[0], connected ? &pl : nullptr, 0, std::max(conn_left, 0)) ||
        !store_tree(root + "/graphr", H[1], G[1], connected ? &pr : nullptr, std::max(conn_right, 0), 0) ||
        !write_bin(root + "/connection.bin", &conn, 1))
      return GPRM_E_INVALID;
  }
  return GPRM_OK;
}

// Radius neighbours on the device (get_near_vertices, prmsolver.cpp:2532-2574,
// over all nodes instead of the q[0] blocks): every (query j, node l) with
// |node_l - query_j|^2 < D^2, summed over the coordinates in order.
extern "C" int gprm_radius_neighbours(int ndof, const float* d_nodes, int n, long long noff, const float* d_queries,
                                      int nq, long long qoff, const float* mins, const float* maxs, float D,
                                      int* d_out_q, int* d_out_node, float* d_out_d2, int cap, long long* total,
                                      void* stream) {
  if (ndof < 1 || ndof > kMaxDofDev || n < 0 || nq < 0 || cap < 0 || !(D > 0.0f) || !mins || !maxs || !total ||
      (n > 0 && (!d_nodes || noff < n)) || (nq > 0 && (!d_queries || qoff < nq)) ||
      (cap > 0 && (!d_out_q || !d_out_node || !d_out_d2)))
    return GPRM_E_INVALID;
  for (int i = 0; i < ndof; ++i)
    if (!(maxs[i] >= mins[i])) return GPRM_E_INVALID;
  *total = 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const CellGrid g = make_cell_grid(ndof, mins, maxs, D);
  DevArr cell_id, counts, starts, cursor, sorted, sid, cnt;
  if (!counts.ensure(sizeof(int) * (size_t)g.ncells) || !starts.ensure(sizeof(int) * (size_t)g.ncells) ||
      !cursor.ensure(sizeof(int) * (size_t)g.ncells) || !cell_id.ensure(sizeof(int) * (size_t)std::max(n, 1)) ||
      !sid.ensure(sizeof(int) * (size_t)std::max(n, 1)) ||
      !sorted.ensure(sizeof(float) * ndof * (size_t)std::max(n, 1)) || !cnt.ensure(sizeof(int)))
    return GPRM_E_HIP;
  int h = 0;
  if (launch_fill_i32(cnt.as<int>(), 0, 1, st) != hipSuccess ||
      launch_grid_build(g, ndof, d_nodes, n, noff, cell_id.as<int>(), counts.as<int>(), starts.as<int>(),
                        cursor.as<int>(), sorted.as<float>(), sid.as<int>(), st) != hipSuccess ||
      launch_near_grid(ndof, d_queries, nq, qoff, g, starts.as<int>(), counts.as<int>(), sorted.as<float>(),
                       sid.as<int>(), n, D * D, d_out_q, d_out_node, d_out_d2, cnt.as<int>(), cap, st) != hipSuccess ||
      hipMemcpyAsync(&h, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GPRM_E_HIP;
  *total = h;
  return GPRM_OK;
}
