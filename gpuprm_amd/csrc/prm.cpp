// prm.cpp -- single-process bidirectional PRM on one GPU.
//
// The planner of the reference (PRMSolver<ndof>, prm/src/lib/prm/prmsolver.
// {h,cpp}) restated for one process and one device: its MPI rank split
// (prmsolver.cpp:1765-1895) becomes one batched step per iteration.  Per
// step (processing_step_part1a..3a, prmsolver.cpp:1973-2187):
//   1. sample num/2 new nodes around random nodes of each tree, accepting a
//      base node m with probability 1/(1+surrnum[m])^3 and drawing uniformly
//      in the cube of half-width D around it (get_random_nodes, :2584-2624);
//      candidates are decided by the exact point check in one batch instead
//      of one host call each;
//   2. list every stored node of either tree closer than D (L2) to each new
//      node (find_neighbours / get_near_vertices, :2532-2574, :2701-2748) with
//      a GPU radius search over a cell index of the first three joint
//      coordinates, in place of the reference's block index over q[0];
//   3. check the candidate edges with the exact edge check (indicator2): all
//      of them, as the reference does, or at most edge_budget per (new node,
//      tree), ranked by the collision-probability model when one is given and
//      by length otherwise;
//   4. insert each new node into every tree it has a free edge to, with all
//      its free checked edges (calc_conn, part3a); a node joining both trees
//      connects them, and Dijkstra on each tree (do_dijkstra, :2804-2846)
//      gives the path.
// The optional model (surrogate.hip) ranks; it never decides.  Sampling bias
// drops candidates before their exact check, edge ranking chooses which
// edges are checked; every node and edge that enters a tree has been decided
// free by the exact checks above.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "../../include/gprm.h"
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
constexpr int kScorePoints = 8;

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

// One tree of the bidirectional search (the reference's `graph`,
// prmsolver.h:40-58, without the block index).
struct Tree {
  int ndof = 0, cap = 0, n = 0;
  int ndev = 0;                         // nodes [0, ndev) are on the device
  std::vector<float> q;                 // AoS ndof*n
  std::vector<int> surrnum;             // neighbours per node
  // adjacency as directed entries in insertion order (both directions of an
  // edge); csr() groups them by node, keeping each node's order
  std::vector<int> ef, et;
  std::vector<float> ew;
  DevArr dq;                            // SoA ndof x cap on the device
  std::vector<float> stage;             // host staging of one upload
  // cell index of the device nodes for the radius search (rebuilt on upload)
  DevArr cell_id, counts, starts, cursor, sorted, sid;
};

int insert(Tree& t, const float* q) {
  if (t.n >= t.cap) return -1;
  int pos = t.n++;
  t.q.insert(t.q.end(), q, q + t.ndof);
  t.surrnum.push_back(0);
  return pos;
}

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

// Upload nodes [ndev, n) to the device copy, one copy per component.
bool flush(Tree& t, hipStream_t st) {
  const int cnt = t.n - t.ndev;
  if (cnt <= 0) return true;
  // the previous upload may still read the staging buffer
  if (hipStreamSynchronize(st) != hipSuccess) return false;
  t.stage.resize((size_t)t.ndof * cnt);
  for (int m = 0; m < cnt; ++m)
    for (int i = 0; i < t.ndof; ++i) t.stage[(size_t)cnt * i + m] = t.q[(size_t)t.ndof * (t.ndev + m) + i];
  for (int i = 0; i < t.ndof; ++i)
    if (hipMemcpyAsync(t.dq.as<float>() + (size_t)t.cap * i + t.ndev, t.stage.data() + (size_t)cnt * i,
                       sizeof(float) * cnt, hipMemcpyHostToDevice, st) != hipSuccess)
      return false;
  t.ndev = t.n;
  return true;
}

// Rebuild the tree's cell index over its nodes [0, ndev).
bool index_nodes(Tree& t, const CellGrid& g, hipStream_t st) {
  if (!t.counts.ensure(sizeof(int) * (size_t)g.ncells) || !t.starts.ensure(sizeof(int) * (size_t)g.ncells) ||
      !t.cursor.ensure(sizeof(int) * (size_t)g.ncells) || !t.cell_id.ensure(sizeof(int) * (size_t)t.cap) ||
      !t.sid.ensure(sizeof(int) * (size_t)t.cap) || !t.sorted.ensure(sizeof(float) * t.ndof * (size_t)t.cap))
    return false;
  return launch_grid_build(g, t.ndof, t.dq.as<float>(), t.ndev, t.cap, t.cell_id.as<int>(), t.counts.as<int>(),
                           t.starts.as<int>(), t.cursor.as<int>(), t.sorted.as<float>(), t.sid.as<int>(),
                           st) == hipSuccess;
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

}  // namespace
}  // namespace gprm

using namespace gprm;

extern "C" int gprm_prm_solve(gprm_ctx* ctx, const float* qstart, const float* qgoal, const gprm_prm_params* prm,
                              gprm_prm_result* res, float* path, int path_cap) {
  if (!ctx || !qstart || !qgoal || !prm || !res) return GPRM_E_INVALID;
  std::memset(res, 0, sizeof(*res));
  const int nd = gprm_dim(ctx);
  if (prm->num < 2 || !(prm->D > 0.0f) || prm->max_nodes < 2 || prm->max_steps < 1 || prm->nbuf < 1 ||
      prm->edge_budget < 0)
    return GPRM_E_INVALID;
  const auto t0 = Clock::now();
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return GPRM_E_HIP;
  struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } } sg{st};
  Surrogate* model = static_cast<Surrogate*>(prm->surrogate);
  const bool bias = model && prm->bias_sampling;
  const bool rank_by_model = model && prm->order_edges;

  Tree T[2];
  for (int g = 0; g < 2; ++g) {
    T[g].ndof = nd;
    T[g].cap = prm->max_nodes;
    if (!T[g].dq.ensure(sizeof(float) * nd * (size_t)T[g].cap)) return GPRM_E_HIP;
  }
  // start and goal must be free (prm/src/main.cpp:135-138)
  int c0 = 0, c1 = 0;
  int rc = gprm_indicator1(ctx, qstart, &c0);
  if (rc == GPRM_OK) rc = gprm_indicator1(ctx, qgoal, &c1);
  if (rc != GPRM_OK) return rc;
  res->point_checks += 2;
  if (c0 != 0 || c1 != 0) return GPRM_OK;  // found = 0
  insert(T[0], qstart);
  insert(T[1], qgoal);
  std::vector<float> lo(nd), hi(nd);
  for (int i = 0; i < nd; ++i) {
    lo[i] = gprm_min(ctx, i);
    hi[i] = gprm_max(ctx, i);
  }
  const CellGrid grid = make_cell_grid(nd, lo.data(), hi.data(), prm->D);
  for (int g = 0; g < 2; ++g)
    if (!flush(T[g], st) || !index_nodes(T[g], grid, st)) return GPRM_E_HIP;

  Rng rng(prm->seed * 0x9E3779B97F4A7C15ULL + 12345);
  const int num = prm->num;
  const float D = prm->D, D2 = prm->D * prm->D;
  const int batch = 4 * num;  // candidates drawn per exact point-check batch
  std::vector<float> keptT;
  std::vector<int> cand_tree(batch), kept, ccode(batch);
  std::vector<float> qnew((size_t)nd * num);
  std::vector<float> prob;
  std::vector<double> sorted_u;
  std::vector<int> slot, base(batch);
  std::vector<size_t> bucket;
  DevArr d_cand, d_code, d_new, d_nq, d_nn, d_nd2, d_cnt, d_qs, d_qe, d_er, d_mq, d_mp, d_eidx, d_base, d_kidx,
      d_kept;
  std::vector<int> heidx;
  if (!d_cand.ensure(sizeof(float) * nd * batch) || !d_code.ensure(sizeof(int) * batch) ||
      !d_new.ensure(sizeof(float) * nd * num) || !d_nq.ensure(sizeof(int) * 2 * (size_t)prm->nbuf) ||
      !d_nn.ensure(sizeof(int) * 2 * (size_t)prm->nbuf) || !d_nd2.ensure(sizeof(float) * 2 * (size_t)prm->nbuf) ||
      !d_cnt.ensure(8) || !d_base.ensure(sizeof(int) * batch) ||
      (bias && (!d_mp.ensure(sizeof(float) * batch) || !d_kidx.ensure(sizeof(int) * batch) ||
                !d_kept.ensure(sizeof(float) * nd * batch))))
    return GPRM_E_HIP;
  std::vector<int> hq, hn, her;
  std::vector<float> hd2, score;

  int conn_left = -1, conn_right = -1;
  for (int step = 0; step < prm->max_steps && conn_left < 0; ++step) {
    res->steps = step + 1;
    // ---- 1. sampling ---------------------------------------------------------
    auto ts = Clock::now();
    int have[2] = {0, 0};
    const int want[2] = {num / 2, num - num / 2};
    int filled = 0;
    int guard = 0;
    // Base nodes: the reference draws a node uniformly and keeps it with
    // probability 1/(1+surrnum)^3, i.e. it samples nodes with weight
    // 1/(1+surrnum)^3; drawn here directly from cumulative weights.
    std::vector<double> cum[2];
    for (int g = 0; g < 2; ++g) {
      cum[g].resize(T[g].n);
      double acc = 0.0;
      for (int m = 0; m < T[g].n; ++m) {
        double x = 1.0 + T[g].surrnum[m];
        acc += 1.0 / (x * x * x);
        cum[g][m] = acc;
      }
    }
    while (filled < num) {
      if (++guard > 1000) break;
      // Candidates alternate between the trees (k & 1; a tree's surplus is
      // discarded).  A tree's base nodes are drawn as sorted variates of its
      // weight distribution: sorted uniforms from exponential spacings, then
      // one merge pass over the cumulative weights (the inverse-CDF draw of
      // the reference's rule, without a search per draw).  Their candidates
      // go to a random permutation of the tree's slots, so that the order of
      // the batch carries no bias.
      for (int g = 0; g < 2; ++g) {
        const Tree& t = T[g];
        const int kg = (batch - g + 1) / 2;  // slots k = g, g + 2, ...
        sorted_u.resize(kg);
        double acc = 0.0;
        for (int j = 0; j < kg; ++j) {
          acc -= std::log1p(-rng.uniform());
          sorted_u[j] = acc;
        }
        acc -= std::log1p(-rng.uniform());
        const double scale = cum[g].back() / acc;
        slot.resize(kg);
        for (int j = 0; j < kg; ++j) slot[j] = g + 2 * j;
        for (int j = kg - 1; j > 0; --j) std::swap(slot[j], slot[rng.integer(0, j)]);
        int m = 0;
        for (int j = 0; j < kg; ++j) {
          const double u = sorted_u[j] * scale;
          while (m < t.n - 1 && cum[g][m] <= u) ++m;
          const int k = slot[j];
          cand_tree[k] = g;
          base[k] = m;
        }
      }
      // the candidates themselves are formed on the device from the trees'
      // device copies (offsets from a counter-based generator keyed per batch)
      if (hipMemcpyAsync(d_base.p, base.data(), sizeof(int) * batch, hipMemcpyHostToDevice, st) != hipSuccess ||
          launch_sample_candidates(nd, batch, d_base.as<int>(), T[0].dq.as<float>(), T[1].dq.as<float>(), T[0].cap,
                                   D, rng.next(), d_cand.as<float>(), st) != hipSuccess)
        return GPRM_E_HIP;
      // Sampling bias: every second pair of candidates must be predicted
      // free; candidates it drops are never checked.  The others, and all
      // candidates without a model, go to the exact check.
      kept.clear();
      const float* d_chk = d_cand.as<float>();
      if (bias) {
        auto tm = Clock::now();
        prob.resize(batch);
        if (surrogate_predict(model, d_cand.as<float>(), batch, batch, d_mp.as<float>(), st) != 0 ||
            hipMemcpyAsync(prob.data(), d_mp.p, sizeof(float) * batch, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
          return GPRM_E_HIP;
        res->model_predictions += batch;
        res->ms_model += ms_since(tm);
        for (int k = 0; k < batch; ++k)
          if (!((k & 2) && prob[k] > 0.5f)) kept.push_back(k);
        const int nk = (int)kept.size();
        if (nk < batch) {
          if (hipMemcpyAsync(d_kidx.p, kept.data(), sizeof(int) * nk, hipMemcpyHostToDevice, st) != hipSuccess ||
              launch_gather_columns(nd, nk, d_kidx.as<int>(), d_cand.as<float>(), batch, d_kept.as<float>(), st) !=
                  hipSuccess)
            return GPRM_E_HIP;
          d_chk = d_kept.as<float>();
        }
      } else {
        kept.resize(batch);
        std::iota(kept.begin(), kept.end(), 0);
      }
      const int nk = (int)kept.size();
      res->candidates_skipped += batch - nk;
      if (nk == 0) continue;
      rc = gprm_indicator_dev(ctx, d_chk, d_code.as<int>(), nk, nk, st);
      if (rc != GPRM_OK) return rc;
      keptT.resize((size_t)nd * nk);
      if (hipMemcpyAsync(ccode.data(), d_code.p, sizeof(int) * nk, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(keptT.data(), d_chk, sizeof(float) * nd * nk, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return GPRM_E_HIP;
      res->point_checks += nk;
      for (int a = 0; a < nk && filled < num; ++a) {
        const int g = cand_tree[kept[a]];
        if (ccode[a] != 0 || have[g] >= want[g]) continue;
        for (int i = 0; i < nd; ++i) qnew[(size_t)nd * filled + i] = keptT[(size_t)nk * i + a];
        ++have[g];
        ++filled;
      }
    }
    res->ms_sampling += ms_since(ts);
    if (filled < num) break;

    // ---- 2. radius neighbours in both trees ---------------------------------
    auto tn = Clock::now();
    std::vector<float> qnewT((size_t)nd * num);
    for (int j = 0; j < num; ++j)
      for (int i = 0; i < nd; ++i) qnewT[(size_t)num * i + j] = qnew[(size_t)nd * j + i];
    if (hipMemcpyAsync(d_new.p, qnewT.data(), sizeof(float) * nd * num, hipMemcpyHostToDevice, st) != hipSuccess)
      return GPRM_E_HIP;
    // candidate edges (new node j, tree g, stored node, distance) in the
    // reference's buffer order: by new node, left tree before right tree
    // (find_neighbours), then by node; at most nbuf per tree and per step.
    // The two trees' searches run back to back into the two halves of the
    // output arrays and are read back together.
    struct Cand { int j, g, node; float d2; };
    std::vector<Cand> cands;
    const size_t half = (size_t)prm->nbuf;
    int cnt[2] = {0, 0};
    {
      const int zero[2] = {0, 0};
      if (hipMemcpyAsync(d_cnt.p, zero, 8, hipMemcpyHostToDevice, st) != hipSuccess) return GPRM_E_HIP;
      for (int g = 0; g < 2; ++g)
        if (launch_near_grid(nd, d_new.as<float>(), num, num, grid, T[g].starts.as<int>(), T[g].counts.as<int>(),
                             T[g].sorted.as<float>(), T[g].sid.as<int>(), T[g].ndev, D2, d_nq.as<int>() + g * half,
                             d_nn.as<int>() + g * half, d_nd2.as<float>() + g * half, d_cnt.as<int>() + g,
                             prm->nbuf, st) != hipSuccess)
          return GPRM_E_HIP;
      if (hipMemcpyAsync(cnt, d_cnt.p, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return GPRM_E_HIP;
      hq.resize(2 * half); hn.resize(2 * half); hd2.resize(2 * half);
      for (int g = 0; g < 2; ++g) {
        cnt[g] = std::min(cnt[g], prm->nbuf);
        if (cnt[g] == 0) continue;
        if (hipMemcpyAsync(hq.data() + g * half, d_nq.as<int>() + g * half, 4 * (size_t)cnt[g],
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hn.data() + g * half, d_nn.as<int>() + g * half, 4 * (size_t)cnt[g],
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hd2.data() + g * half, d_nd2.as<float>() + g * half, 4 * (size_t)cnt[g],
                           hipMemcpyDeviceToHost, st) != hipSuccess)
          return GPRM_E_HIP;
      }
      if (hipStreamSynchronize(st) != hipSuccess) return GPRM_E_HIP;
    }
    // counting sort by (j, g), then each (small) group by node
    bucket.assign(2 * (size_t)num + 1, 0);
    for (int g = 0; g < 2; ++g)
      for (int i = 0; i < cnt[g]; ++i) ++bucket[2 * (size_t)hq[g * half + i] + g + 1];
    for (size_t b = 1; b < bucket.size(); ++b) bucket[b] += bucket[b - 1];
    cands.resize((size_t)cnt[0] + cnt[1]);
    for (int g = 0; g < 2; ++g)
      for (int i = 0; i < cnt[g]; ++i) {
        const size_t x = g * half + i;
        cands[bucket[2 * (size_t)hq[x] + g]++] = Cand{hq[x], g, hn[x], hd2[x]};
      }
    for (size_t a = 0, b = 0; a < cands.size(); a = b) {
      b = a + 1;
      while (b < cands.size() && cands[b].j == cands[a].j && cands[b].g == cands[a].g) ++b;
      std::sort(cands.begin() + a, cands.begin() + b, [](const Cand& x, const Cand& y) { return x.node < y.node; });
    }
    if ((long long)cands.size() > prm->nbuf) cands.resize(prm->nbuf);
    res->ms_near += ms_since(tn);

    // ---- 2b. edge budget: rank the candidates of each (new node, tree) -------
    if (prm->edge_budget > 0) {
      const int B = prm->edge_budget;
      // groups of equal (j, g) are contiguous
      std::vector<size_t> gstart;
      for (size_t i = 0; i < cands.size(); ++i)
        if (i == 0 || cands[i].j != cands[i - 1].j || cands[i].g != cands[i - 1].g) gstart.push_back(i);
      gstart.push_back(cands.size());
      // the model scores only edges of groups with more than B candidates
      score.assign(cands.size(), 0.0f);
      if (rank_by_model) {
        auto tm = Clock::now();
        std::vector<size_t> need;
        for (size_t a = 0; a + 1 < gstart.size(); ++a)
          if (gstart[a + 1] - gstart[a] > (size_t)B)
            for (size_t i = gstart[a]; i < gstart[a + 1]; ++i) need.push_back(i);
        const int nn = (int)need.size();
        const int np = nn * kScorePoints;
        // the scored points are formed on the device from the new nodes and
        // the trees' device copies; only the indices travel
        heidx.resize(3 * (size_t)nn);
        for (int a = 0; a < nn; ++a) {
          const Cand& c = cands[need[a]];
          heidx[a] = c.j;
          heidx[nn + a] = c.g;
          heidx[2 * (size_t)nn + a] = c.node;
        }
        prob.resize(np);
        if (np > 0) {
          if (!d_eidx.ensure(sizeof(int) * 3 * (size_t)nn) || !d_mq.ensure(sizeof(float) * nd * (size_t)np) ||
              !d_mp.ensure(sizeof(float) * (size_t)np))
            return GPRM_E_HIP;
          if (hipMemcpyAsync(d_eidx.p, heidx.data(), sizeof(int) * 3 * (size_t)nn, hipMemcpyHostToDevice, st) !=
                  hipSuccess ||
              launch_score_points(nd, nn, kScorePoints, d_eidx.as<int>(), d_new.as<float>(), num, T[0].dq.as<float>(),
                                  T[1].dq.as<float>(), T[0].cap, d_mq.as<float>(), st) != hipSuccess ||
              surrogate_predict(model, d_mq.as<float>(), np, np, d_mp.as<float>(), st) != 0 ||
              hipMemcpyAsync(prob.data(), d_mp.p, sizeof(float) * (size_t)np, hipMemcpyDeviceToHost, st) !=
                  hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess)
            return GPRM_E_HIP;
        }
        res->model_predictions += np;
        for (size_t a = 0; a < need.size(); ++a) {
          float m = 0.0f;
          for (int s = 0; s < kScorePoints; ++s) m = std::max(m, prob[a * kScorePoints + s]);
          score[need[a]] = m;
        }
        res->ms_model += ms_since(tm);
      }
      std::vector<Cand> keep;
      std::vector<size_t> idx;
      for (size_t a = 0; a + 1 < gstart.size(); ++a) {
        const size_t b0 = gstart[a], b1 = gstart[a + 1];
        idx.resize(b1 - b0);
        std::iota(idx.begin(), idx.end(), b0);
        if (b1 - b0 > (size_t)B) {
          std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
            return score[x] != score[y] ? score[x] < score[y] : cands[x].d2 < cands[y].d2;
          });
          idx.resize(B);
          std::sort(idx.begin(), idx.end());
        }
        for (size_t i : idx) keep.push_back(cands[i]);
      }
      res->edges_skipped += (long long)(cands.size() - keep.size());
      cands.swap(keep);
    }
    std::vector<int> eq(cands.size()), eg(cands.size()), en(cands.size());
    std::vector<float> ed(cands.size());
    for (size_t i = 0; i < cands.size(); ++i) {
      eq[i] = cands[i].j;
      eg[i] = cands[i].g;
      en[i] = cands[i].node;
      ed[i] = std::sqrt(cands[i].d2);
    }

    // ---- 3. exact edge checks ------------------------------------------------
    auto te = Clock::now();
    const int E = (int)eq.size();
    her.assign(E, 1);
    if (E > 0) {
      if (!d_qs.ensure(sizeof(float) * nd * (size_t)E) || !d_qe.ensure(sizeof(float) * nd * (size_t)E) ||
          !d_er.ensure(sizeof(int) * (size_t)E) || !d_eidx.ensure(sizeof(int) * 3 * (size_t)E))
        return GPRM_E_HIP;
      // endpoints gathered on the device from the new nodes (d_new) and the
      // trees' device copies; only the indices travel
      heidx.resize(3 * (size_t)E);
      std::copy(eq.begin(), eq.end(), heidx.begin());
      std::copy(eg.begin(), eg.end(), heidx.begin() + E);
      std::copy(en.begin(), en.end(), heidx.begin() + 2 * (size_t)E);
      if (hipMemcpyAsync(d_eidx.p, heidx.data(), sizeof(int) * 3 * (size_t)E, hipMemcpyHostToDevice, st) !=
              hipSuccess ||
          launch_gather_edges(nd, E, d_eidx.as<int>(), d_new.as<float>(), num, T[0].dq.as<float>(),
                              T[1].dq.as<float>(), T[0].cap, d_qs.as<float>(), d_qe.as<float>(), st) != hipSuccess)
        return GPRM_E_HIP;
      gprm_edge_stats es;
      rc = gprm_indicator2_dev(ctx, d_qs.as<float>(), d_qe.as<float>(), d_er.as<int>(), E, E, st, &es);
      if (rc != GPRM_OK) return rc;
      if (hipMemcpyAsync(her.data(), d_er.p, sizeof(int) * E, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return GPRM_E_HIP;
      res->edges_checked += E;
      res->edge_samples += es.samples_total;
    }
    res->ms_edges += ms_since(te);

    // ---- 4. insert nodes with free edges; connection ------------------------
    auto th = Clock::now();
    size_t e = 0;
    for (int j = 0; j < num && conn_left < 0; ++j) {
      int pos[2] = {-1, -1};
      while (e < (size_t)E && eq[e] < j) ++e;
      const size_t e0 = e;
      size_t e1 = e0;
      while (e1 < (size_t)E && eq[e1] == j) ++e1;
      for (int g = 0; g < 2; ++g) {
        bool any = false;
        for (size_t k = e0; k < e1; ++k)
          if (eg[k] == g && her[k] == 0) { any = true; break; }
        if (!any) continue;
        pos[g] = insert(T[g], &qnew[(size_t)nd * j]);
        if (pos[g] < 0) break;
        for (size_t k = e0; k < e1; ++k) {
          if (eg[k] != g || her[k] != 0) continue;
          const int other = en[k];
          Tree& tg = T[g];
          tg.ef.push_back(pos[g]);
          tg.et.push_back(other);
          tg.ew.push_back(ed[k]);
          tg.ef.push_back(other);
          tg.et.push_back(pos[g]);
          tg.ew.push_back(ed[k]);
          ++T[g].surrnum[other];
          ++T[g].surrnum[pos[g]];
          ++res->edges_free;
        }
      }
      e = e1;
      if (pos[0] >= 0 && pos[1] >= 0) {
        conn_left = pos[0];
        conn_right = pos[1];
      }
      if ((pos[0] < 0 && T[0].n >= T[0].cap) || (pos[1] < 0 && T[1].n >= T[1].cap)) break;
    }
    for (int g = 0; g < 2; ++g)
      if (!flush(T[g], st) || !index_nodes(T[g], grid, st)) return GPRM_E_HIP;
    res->ms_host += ms_since(th);
    if (T[0].n >= T[0].cap || T[1].n >= T[1].cap) break;
  }
  res->nodes_left = T[0].n;
  res->nodes_right = T[1].n;
  std::vector<int> pl, pr;
  bool connected = false;
  Csr G[2];
  if (conn_left >= 0 || prm->output_dir)
    for (int g = 0; g < 2; ++g) G[g] = csr(T[g]);
  if (conn_left >= 0) {
    connected = dijkstra(T[0], G[0], 0, conn_left, &pl) && dijkstra(T[1], G[1], conn_right, 0, &pr);
    if (connected) {
      // pl = [conn_left, ..., start], pr = [goal, ..., conn_right]; conn_left
      // and conn_right hold the same configuration
      std::vector<const float*> seq;
      for (int i = (int)pl.size() - 1; i >= 0; --i) seq.push_back(&T[0].q[(size_t)nd * pl[i]]);
      for (int i = (int)pr.size() - 2; i >= 0; --i) seq.push_back(&T[1].q[(size_t)nd * pr[i]]);
      res->found = 1;
      res->path_len = (int)seq.size();
      if (path)
        for (int i = 0; i < std::min(path_cap, res->path_len); ++i) std::memcpy(path + (size_t)nd * i, seq[i], 4 * nd);
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess) return GPRM_E_HIP;
  res->ms_total = ms_since(t0);
  if (prm->output_dir) {
    // store_results (prmsolver.cpp:2947-2973): graphl from the start root to
    // the connection, graphr from the connection to the goal root
    const std::string root(prm->output_dir);
    const int conn = connected ? 1 : 0;
    std::error_code ec;
    std::filesystem::create_directories(root, ec);
    if (ec || !store_tree(root + "/graphl", T[0], G[0], connected ? &pl : nullptr, 0, std::max(conn_left, 0)) ||
        !store_tree(root + "/graphr", T[1], G[1], connected ? &pr : nullptr, std::max(conn_right, 0), 0) ||
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
