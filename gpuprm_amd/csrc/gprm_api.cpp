// gprm_api.cpp -- host runtime behind the C ABI of include/gprm.h.
//
// One context = one scene on one device (the reference's RobotConfigspace,
// prm/src/lib/robotspace/robotconfigspace.{h,cpp}).  The scene is packed once
// into a device image that the separating-vector kernel stages in LDS.  Work
// is split into chunks of at most kItemBudget configurations, each chunk
// being one kinematics launch and one persistent separating-vector launch.
// Edge queries run in rounds of growing sample windows; after every round the
// samples of edges that are already blocked are no longer scheduled.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gprm.h"
#include "context.h"
#include "glibc_trig.h"
#include "runtime.h"
#include "scene.h"

namespace gprm {
namespace {

// Configurations per kinematics / separating-vector launch: 2^24 (frames of
// up to 5.4 GB for a 7-DOF arm; HBM holds 288 GB).  On the 6-DOF edge
// benchmark 2^24 runs each window in one launch and measured 5.38 against
// 5.47 ms per step for 2^23.
constexpr int kItemBudget = 1 << 24;
// Sample positions in the first edge round; each later round is kWindowGrowth
// times wider.  The golden-ratio order spreads the first positions along the
// edge, so blocked edges are mostly found in the first round.  On the 6-DOF
// edge benchmark, with doubling windows, a first window of 2/4/8/16/32/64/128
// gave 6.80/6.70/6.53/6.37-6.42/6.45/6.57/6.72 ms per step; later, (first
// window, growth) of (16, 2)/(16, 3)/(16, 4)/(16, 5)/(8, 3)/(12, 3)/(24, 3)/
// (32, 3)/(24, 4) gave 5.76/5.56-5.61/5.61-5.62/5.61/5.77/5.90/5.62/5.69/
// 5.82 ms.  Once a short remainder joined the last window and the frame
// planes were 1 KB-aligned, (16, 3)/(16, 4)/(16, 5)/(16, 6)/(8, 4)/(12, 4)/
// (20, 4)/(24, 3)/(32, 3) gave 5.31/5.20/5.27/5.28/5.45/5.26-5.29/5.28/
// 5.27-5.29/5.29 ms.  Once the later rounds' kinematics overlapped the
// previous round, (12, 4)/(16, 4)/(10, 4)/(8, 4)/(16, 3)/(12, 3)/(8, 3)/
// (16, 5)/(24, 4) gave 4.84-4.88/4.91-4.92/4.93-4.94/5.11/4.89/5.09-5.11/
// 5.09-5.10/4.93-4.95/5.01 ms: windows [0,12), [12,60), [60,209) for its
// longest edges.
constexpr int kFirstWindow = 12;
constexpr int kWindowGrowth = 4;
// A call whose whole plan is known (the planner's, bounded by D / dq) and
// holds at most this many pair tests (edges x largest testnum x pairs) runs
// as one round.  Each round costs a launch whose end waits for its longest
// full test (tens of microseconds), and on the planner's steps almost every
// candidate edge is free, so the early exit between rounds saves little.
// Median time-to-path with three rounds / one round up to 2^28 / 2^30 /
// 2^32 pair tests: config1 at 32 nodes per step 226 / 162 / 163 / 162 ms,
// 7-DOF 153 / 151 / 140 / 140 ms, 6-DOF at 1,024 nodes 8.1 / 5.7-5.8 /
// 5.6 / 5.8 ms, with an edge budget at 4,096 nodes 5.7 / 5.8 / 5.1 / 4.9 ms.
constexpr long long kOneRoundPairTests = 1LL << 32;

thread_local std::string g_create_error;

}  // namespace

// Sample-window rounds: positions [0,w0), then widths growing kWindowGrowth-fold;
// one window when pair_tests (a bound on the call's, or -1) is small.
std::vector<std::pair<int, int>> edge_windows(int maxT, long long pair_tests) {
  std::vector<std::pair<int, int>> w;
  if (pair_tests >= 0 && pair_tests <= kOneRoundPairTests && maxT > 0) {
    w.push_back({0, maxT});
    return w;
  }
  int lo = 0, width = kFirstWindow;
  while (lo < maxT) {
    int hi = (int)std::min<long long>((long long)lo + width, maxT);
    // A remainder under a quarter of this window joins it: a round of its
    // own costs about 50 us (four launches and their gaps) however few
    // samples it holds.
    if ((long long)maxT - hi < width / 4) hi = maxT;
    w.push_back({lo, hi});
    lo = hi;
    if (width < (1 << 20)) width *= kWindowGrowth;
  }
  return w;
}

}  // namespace gprm

namespace gprm {
namespace {

int fail(gprm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_fail(gprm_ctx* c, hipError_t e, const char* what) {
  return fail(c, GPRM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define GPRM_HIP(call, what)                      \
  do {                                            \
    hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) return hip_fail(c, e_, what); \
  } while (0)

void fill_robot(const Scene& s, DevRobot* rb) {
  std::memset(rb, 0, sizeof(*rb));
  rb->ndof = s.ndof;
  for (int i = 0; i < s.ndof; ++i) {
    rb->a[i] = s.a[i]; rb->alpha[i] = s.alpha[i]; rb->q[i] = s.q[i]; rb->d[i] = s.d[i];
    rb->mins[i] = s.mins[i]; rb->maxs[i] = s.maxs[i]; rb->types[i] = s.types[i];
    rb->ca[i] = gprm_trig::cosf_ref(s.alpha[i]);
    rb->sa[i] = gprm_trig::sinf_ref(s.alpha[i]);
  }
  rb->c0 = gprm_trig::cosf_ref(0.0f);
  rb->s0 = gprm_trig::sinf_ref(0.0f);
}

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Pack the scene image (layout documented at SceneImg).  The kernel visits
// pairs grouped by the frames they need; each pair keeps its index in the
// reference's CRS order (polytope4data.h:343-357: by `from`, then as given),
// which decides which colliding pair's code a configuration returns.
int build_image(gprm_ctx* c, std::vector<char>* out) {
  const Scene& s = c->scene;
  const int N = (int)s.polys.size();
  std::vector<int> vbase(N);
  int sumn = 0;
  for (int k = 0; k < N; ++k) {
    vbase[k] = sumn;
    sumn += s.polys[k].n();
  }
  if (sumn >= 65536) return fail(c, GPRM_E_CAPACITY, "scene has 65536 or more vertices");
  if (N >= 65536) return fail(c, GPRM_E_CAPACITY, "scene has 65536 or more polytopes");
  // adjacency rows padded to multiples of 4 entries
  std::vector<int> row_off(sumn);
  size_t nadj = 0;
  for (int k = 0; k < N; ++k)
    for (int i = 0; i < s.polys[k].n(); ++i) {
      if (s.polys[k].cnt[i] > 255) return fail(c, GPRM_E_CAPACITY, "vertex degree above 255");
      row_off[vbase[k] + i] = (int)nadj;
      nadj += (size_t)((s.polys[k].cnt[i] + 3) / 4) * 4;
    }
  if (nadj >= (1u << 23)) return fail(c, GPRM_E_CAPACITY, "scene adjacency too large");
  const int M = (int)s.from.size();
  SceneImg& img = c->img;
  img.sumn = sumn; img.summ = (int)nadj; img.M = M; img.N = N;
  size_t off = 0;
  img.off_verts = (int)off; off = align16(off + 16 * (size_t)sumn);
  img.off_pairs = (int)off; off = align16(off + 16 * (size_t)M);
  img.off_adj = (int)off;   off = align16(off + 2 * nadj);
  img.off_spheres = (int)off; off = align16(off + 32 * (size_t)N);
  img.bytes = (int)off;
  out->assign(off, 0);
  char* b = out->data();
  for (int k = 0; k < N; ++k) {
    const Polytope& p = s.polys[k];
    for (int i = 0; i < p.n(); ++i) {
      const int g = vbase[k] + i;
      int info = (row_off[g] << 8) | p.cnt[i];
      float v[4] = {p.verts[3 * i], p.verts[3 * i + 1], p.verts[3 * i + 2], 0.0f};
      std::memcpy(&v[3], &info, 4);
      std::memcpy(b + img.off_verts + 16 * (size_t)g, v, 16);
      const int padded = (p.cnt[i] + 3) / 4 * 4;
      for (int j = 0; j < padded; ++j) {
        unsigned short id = (unsigned short)(j < p.cnt[i] ? vbase[k] + p.dest[p.dsp[i] + j] : g);
        std::memcpy(b + img.off_adj + 2 * (size_t)(row_off[g] + j), &id, 2);
      }
    }
  }
  // bounds per polytope, in its own frame: the centre c of the vertex
  // bounding box with the largest vertex distance from it (sphere radius),
  // then the largest |v_a - c_a| per axis (box half extents); both from
  // doubles, enlarged by a relative 1e-6 and rounded up
  for (int k = 0; k < N; ++k) {
    const Polytope& p = s.polys[k];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < p.n(); ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], (double)p.verts[3 * i + a]);
        hi[a] = std::max(hi[a], (double)p.verts[3 * i + a]);
      }
    float c[3];
    for (int a = 0; a < 3; ++a) c[a] = (float)(0.5 * (lo[a] + hi[a]));
    double r2 = 0.0;
    for (int i = 0; i < p.n(); ++i) {
      double d2 = 0.0;
      for (int a = 0; a < 3; ++a) {
        double dd = (double)p.verts[3 * i + a] - (double)c[a];
        d2 += dd * dd;
      }
      r2 = std::max(r2, d2);
    }
    auto up = [](double x) {
      x = x * (1.0 + 1e-6) + 1e-30;
      float f = (float)x;
      if ((double)f < x) f = std::nextafter(f, INFINITY);
      return f;
    };
    double h[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < p.n(); ++i)
      for (int a = 0; a < 3; ++a) h[a] = std::max(h[a], std::fabs((double)p.verts[3 * i + a] - (double)c[a]));
    const float rf = up(std::sqrt(r2));
    // w of the second record: |c|_1 + r, the polytope's magnitude term of the
    // bound test's margin in its own frame (the world frame for obstacles)
    const float mag = (float)(std::fabs((double)c[0]) + std::fabs((double)c[1]) + std::fabs((double)c[2]) + rf);
    float sp[8] = {c[0], c[1], c[2], rf, up(h[0]), up(h[1]), up(h[2]), mag};
    std::memcpy(b + img.off_spheres + 32 * (size_t)k, sp, 32);
  }
  std::vector<int> order(M);
  for (int i = 0; i < M; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    int kx = s.sys[s.to[x]] * 256 + s.sys[s.from[x]], ky = s.sys[s.to[y]] * 256 + s.sys[s.from[y]];
    return kx < ky;
  });
  // The kernel keeps the last P frame and the last Q frame of a segment.
  // Within a block of equal Q frame, the pairs whose P is also a moving link
  // (link-link pairs) each load their P frame unless it is the one kept, so
  // each block's P frames are ordered to start with the frame the previous
  // block ended with and, where possible, to end with one the next block
  // needs (on the 6-DOF arm: 7 P loads per item instead of 10).  The order
  // of the pairs does not change results: edge results only ask whether any
  // pair collides, point codes take the smallest CRS index (atomicMin).
  {
    const auto sysP = [&](int i) { return s.sys[s.from[i]]; };
    const auto sysQ = [&](int i) { return s.sys[s.to[i]]; };
    std::vector<std::pair<int, int>> blocks;  // [begin, end) of equal sysQ
    for (int j = 0; j < M;) {
      int e = j;
      while (e < M && sysQ(order[e]) == sysQ(order[j])) ++e;
      blocks.push_back({j, e});
      j = e;
    }
    int prevP = -1;
    for (size_t b = 0; b < blocks.size(); ++b) {
      const int j0 = blocks[b].first, j1 = blocks[b].second;
      // the block's P frames in sorted order (runs of equal sysP stay together)
      std::vector<int> ps;
      for (int j = j0; j < j1; ++j)
        if (sysP(order[j]) != 0 && (ps.empty() || ps.back() != sysP(order[j]))) ps.push_back(sysP(order[j]));
      if (ps.empty()) continue;
      std::vector<int> next;
      if (b + 1 < blocks.size())
        for (int j = blocks[b + 1].first; j < blocks[b + 1].second; ++j)
          if (sysP(order[j]) != 0) next.push_back(sysP(order[j]));
      auto in_next = [&](int v) { return std::find(next.begin(), next.end(), v) != next.end(); };
      // the kept frame first, then the others; the first of those the next
      // block needs goes last
      std::vector<int> seq, rest;
      if (std::find(ps.begin(), ps.end(), prevP) != ps.end()) seq.push_back(prevP);
      for (int v : ps)
        if (v != prevP) rest.push_back(v);
      for (size_t k = 0; k < rest.size(); ++k)
        if (in_next(rest[k])) {
          std::rotate(rest.begin() + k, rest.begin() + k + 1, rest.end());
          break;
        }
      seq.insert(seq.end(), rest.begin(), rest.end());
      // stable regroup: obstacle pairs (sysP 0) first, then the P runs in seq
      std::vector<int> reord;
      for (int j = j0; j < j1; ++j)
        if (sysP(order[j]) == 0) reord.push_back(order[j]);
      for (int v : seq)
        for (int j = j0; j < j1; ++j)
          if (sysP(order[j]) == v) reord.push_back(order[j]);
      std::copy(reord.begin(), reord.end(), order.begin() + j0);
      prevP = seq.back();
    }
  }
  for (int j = 0; j < M; ++j) {
    const int i = order[j];
    int pr[4] = {vbase[s.from[i]] | (vbase[s.to[i]] << 16), s.sys[s.from[i]] | (s.sys[s.to[i]] << 8), i,
                 s.from[i] | (s.to[i] << 16)};
    std::memcpy(b + img.off_pairs + 16 * (size_t)j, pr, 16);
  }
  return GPRM_OK;
}

int new_ctx(Scene&& s, float dq, int nbuf, gprm_ctx** out) {
  std::string err = validate(s);
  if (!err.empty()) return fail(nullptr, GPRM_E_SCENE, err);
  if (!(dq > 0.0f) || !std::isfinite(dq)) return fail(nullptr, GPRM_E_INVALID, "dq must be positive");
  gprm_ctx* c = new gprm_ctx();
  c->scene = std::move(s);
  c->dq = dq;
  c->nbuf = nbuf;
  fill_robot(c->scene, &c->rb);
  *out = c;
  return GPRM_OK;
}

int acquire_slot(gprm_ctx* c, int* idx) {
  for (size_t i = 0; i < c->slots.size(); ++i)
    if (!c->slots[i]->busy) { *idx = (int)i; c->slots[i]->busy = true; return GPRM_OK; }
  std::unique_ptr<Slot> s(new Slot());
  GPRM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
  s->busy = true;
  c->slots.push_back(std::move(s));
  *idx = (int)c->slots.size() - 1;
  return GPRM_OK;
}

// Profiling helpers: bracket a launch with events on its stream.
hipError_t prof_begin(gprm_ctx* c, Slot& sl, int kind, hipStream_t st, size_t* idx) {
  if (!c->profile) return hipSuccess;
  if (sl.ev_used == sl.ev_start.size()) {
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) { (void)hipEventDestroy(a); return e; }
    sl.ev_start.push_back(a);
    sl.ev_end.push_back(b);
    sl.ev_kind.push_back(0);
  }
  *idx = sl.ev_used++;
  sl.ev_kind[*idx] = kind;
  return hipEventRecord(sl.ev_start[*idx], st);
}
hipError_t prof_end(gprm_ctx* c, Slot& sl, size_t idx, hipStream_t st) {
  if (!c->profile) return hipSuccess;
  return hipEventRecord(sl.ev_end[idx], st);
}
unsigned long long* tally_ptr(gprm_ctx* c, Slot& sl) {
  return c->profile ? sl.tally.as<unsigned long long>() : nullptr;
}
// The slot's tally counters are cleared once and then accumulate over calls;
// gprm_profile_read folds them into the context totals, so a profiled call
// pays no extra copy or synchronisation for them.
int prof_prepare(gprm_ctx* c, Slot& sl, hipStream_t st) {
  sl.ev_used = 0;
  if (!c->profile || sl.tally_pending) return GPRM_OK;
  GPRM_HIP(sl.tally.ensure(sizeof(unsigned long long) * kTallyCount * kTallySlots), "hipMalloc tally");
  GPRM_HIP(launch_fill_u32(sl.tally.as<unsigned>(), 0u, 2 * kTallyCount * kTallySlots, st), "fill");
  sl.tally_pending = true;
  return GPRM_OK;
}
// After the slot's stream has completed: fold the launches' event times into
// the context totals.
int prof_harvest(gprm_ctx* c, Slot& sl) {
  for (size_t i = 0; i < sl.ev_used; ++i) {
    float ms = 0.0f;
    GPRM_HIP(hipEventElapsedTime(&ms, sl.ev_start[i], sl.ev_end[i]), "hipEventElapsedTime");
    if (sl.ev_kind[i] == 0) { c->prof.sva_launches++; c->prof.sva_ms += ms; }
    else { c->prof.fk_launches++; c->prof.fk_ms += ms; }
  }
  sl.ev_used = 0;
  return GPRM_OK;
}
// Fold the accumulated tally counters of an idle slot (its stream has
// completed) into the context totals.
int prof_fold_tally(gprm_ctx* c, Slot& sl) {
  if (sl.tally_pending) {
    std::vector<unsigned long long> slots((size_t)kTallyCount * kTallySlots);
    GPRM_HIP(hipMemcpy(slots.data(), sl.tally.p, sizeof(unsigned long long) * slots.size(), hipMemcpyDeviceToHost),
             "copy tally");
    unsigned long long t[kTallyCount] = {};
    for (int k = 0; k < kTallySlots; ++k)
      for (int i = 0; i < kTallyCount; ++i) t[i] += slots[(size_t)k * kTallyCount + i];
    c->prof.pairs_separated += (long long)t[0];
    c->prof.pairs_colliding += (long long)t[1];
    c->prof.units_dropped += (long long)t[2];
    c->prof.units_done += (long long)t[3];
    c->prof.iterations += (long long)t[4];
    c->prof.bound_separated += (long long)t[5];
    sl.tally_pending = false;
  }
  return GPRM_OK;
}

}  // namespace

// Pairs per work unit of a launch over about n items: the shape's, or fewer
// so that there are at least two units (chunk x 64-item group) per resident
// wave; never below 8.
int unit_pairs(const gprm_ctx* c, long long n) {
  const long long waves = (long long)c->shape.grid * (c->shape.block / 64);
  const long long groups = std::max(1LL, (n + 63) / 64);
  long long want = groups * c->img.M / std::max(1LL, 2 * waves);
  want = std::max(8LL, std::min((long long)c->shape.ppu, want));
  return (int)want;
}

// The context's shape with the grid cut to the blocks a launch of `units`
// work units can use (small launches then dispatch and stage fewer blocks).
SvaShape launch_shape(const gprm_ctx* c, long long units) {
  SvaShape sh = c->shape;
  const long long per_block = std::max(1, sh.block / 64);
  const long long need = std::max(1LL, (units + per_block - 1) / per_block);
  if (need < sh.grid) sh.grid = (int)need;
  return sh;
}

namespace {

// Items of one launch: at most kItemBudget, and few enough that the unit
// count stays below 2^31 even with the smallest units (8 pairs).
int item_budget(const gprm_ctx* c) {
  const int upi = (c->img.M + 7) / 8;  // units per item at the smallest unit
  return std::max(1, std::min(kItemBudget, 0x7fffffff / std::max(1, upi)));
}

// Point checks of n items on the device (q SoA with offset, res[n]).
int run_points(gprm_ctx* c, Slot& sl, const float* d_q, long long offset, int n, int* d_res,
               hipStream_t st, const unsigned char* skip = nullptr) {
  if (n <= 0) return GPRM_OK;
  const int nd = c->scene.ndof;
  const int chunk = std::min(n, item_budget(c));
  GPRM_HIP(sl.frames.ensure(sizeof(float) * 12 * nd * (size_t)frame_stride(chunk)), "hipMalloc frames");
  GPRM_HIP(sl.status.ensure(sizeof(int) * (size_t)chunk), "hipMalloc words");
  const int nchunks = (n + chunk - 1) / chunk;
  GPRM_HIP(sl.hist.ensure(sizeof(float) * 4 * kMaxIter * (size_t)c->shape.grid * c->shape.block), "hipMalloc history");
  GPRM_HIP(sl.counters.ensure(sizeof(unsigned) * (size_t)nchunks), "hipMalloc counters");
  GPRM_HIP(sl.groups.ensure(sizeof(int) * (size_t)((chunk + 63) / 64)), "hipMalloc group counters");
  for (int ci = 0; ci < nchunks; ++ci) {
    int c0 = ci * chunk, cn = std::min(chunk, n - c0);
    size_t ev = 0;
    GPRM_HIP(prof_begin(c, sl, 1, st, &ev), "event");
    GPRM_HIP(launch_fk_points(c->rb, d_q + c0, offset, cn, sl.frames.as<float>(), frame_stride(cn),
                              sl.status.as<int>(), 1,
                              sl.counters.as<unsigned>() + ci, st, skip ? skip + c0 : nullptr,
                              sl.groups.as<int>()),
             "kinematics launch");
    GPRM_HIP(prof_end(c, sl, ev, st), "event");
    SvaArgs a{};
    a.img = c->img;
    a.ndof = nd;
    a.frames = sl.frames.as<float>();
    a.fstride = frame_stride(cn);
    a.hist = sl.hist.as<float4>();
    a.nitems = cn;
    a.counter = sl.counters.as<unsigned>() + ci;
    a.pairs_per_unit = unit_pairs(c, cn);
    a.units_per_item = (c->img.M + a.pairs_per_unit - 1) / a.pairs_per_unit;
    a.word = sl.status.as<int>();
    // the separating-vector launch writes the codes (no finalize launch)
    a.group_done = sl.groups.as<int>();
    a.codes = d_res + c0;
    a.tally = tally_ptr(c, sl);
    a.static_first = 1;
    GPRM_HIP(prof_begin(c, sl, 0, st, &ev), "event");
    GPRM_HIP(launch_sva(a, 0, launch_shape(c, (long long)((cn + 63) / 64) * a.units_per_item), st),
             "separating-vector launch");
    GPRM_HIP(prof_end(c, sl, ev, st), "event");
  }
  return GPRM_OK;
}


// Edge checks on the device.  testnum/stride/res are computed by the
// prologue kernel; maxT (host-known or read back) sizes the rounds.
int run_edges(gprm_ctx* c, Slot& sl, const float* d_qs, const float* d_qe, long long offset, int N,
              int* d_res, hipStream_t st, int host_maxT, gprm_edge_stats* stats, int maxT_hint = 0,
              const unsigned char* skip = nullptr) {
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (N <= 0) return GPRM_OK;
  const int nd = c->scene.ndof;
  GPRM_HIP(sl.testnum.ensure(sizeof(int) * (size_t)N), "hipMalloc testnum");
  GPRM_HIP(sl.stride.ensure(sizeof(int) * (size_t)N), "hipMalloc stride");
  GPRM_HIP(sl.misc.ensure(64), "hipMalloc misc");
  int* d_max = reinterpret_cast<int*>(sl.misc.as<char>() + kMiscMaxT);
  unsigned long long* d_total = reinterpret_cast<unsigned long long*>(sl.misc.as<char>() + kMiscTotal);
  GPRM_HIP(launch_fill_u32(sl.misc.as<unsigned>(), 0u, 16, st), "fill");
  EdgeArgs ea{};
  ea.qs = d_qs; ea.qe = d_qe; ea.offset = offset; ea.nedges = N;
  ea.skip = skip;
  ea.testnum = sl.testnum.as<int>(); ea.stride = sl.stride.as<int>(); ea.edge_res = d_res;
  GPRM_HIP(launch_edge_prologue(c->rb, c->dq, ea, sl.testnum.as<int>(), sl.stride.as<int>(), d_max, d_total, st),
           "edge prologue launch");
  // The first window [0, kFirstWindow) needs no largest testnum (edges with
  // fewer samples emit fewer items), so it is enqueued at once; the largest
  // testnum, which sizes the later windows, is read back on the side stream
  // meanwhile.
  const bool need_max = host_maxT < 0 || stats;
  if (need_max) {
    if (!sl.side) GPRM_HIP(hipStreamCreateWithFlags(&sl.side, hipStreamNonBlocking), "hipStreamCreate");
    if (!sl.ev_prologue) GPRM_HIP(hipEventCreateWithFlags(&sl.ev_prologue, hipEventDisableTiming), "hipEventCreate");
    if (!sl.pinned) GPRM_HIP(hipHostMalloc(&sl.pinned, 32, hipHostMallocDefault), "hipHostMalloc");
    GPRM_HIP(hipEventRecord(sl.ev_prologue, st), "event");
  }
  // launches: edges [e0, e1) x order positions [lo, hi), at most item_budget
  // items each
  struct Launch { int e0, e1, lo, hi; };
  std::vector<Launch> plan;
  std::vector<int> early;  // item counts of launches whose counters were replaced
  auto add_window = [&](int lo, int hi) {
    const int per = std::max(1, item_budget(c) / (hi - lo));
    for (int e0 = 0; e0 < N; e0 += per) plan.push_back({e0, std::min(N, e0 + per), lo, hi});
  };
  // Launch i uses buffer set i & 1.  Launch 0 runs on the call's stream;
  // launches i >= 1 compact and compute frames on the prep stream (after
  // launch 0's kinematics, and once launch i - 2's separating-vector kernel
  // has released their buffer set), so that this work overlaps launch
  // i - 1's separating-vector kernel.  Their compaction reads the open edges
  // while launch i - 1 may still block some: it can schedule samples of
  // edges that end up blocked, which the separating-vector kernel skips
  // (its units test the edge's result first) or tests without changing the
  // decision.
  auto set_frames = [&](size_t i) -> DevBuf& { return (i & 1) ? sl.frames_odd : sl.frames; };
  auto set_edge = [&](size_t i) -> DevBuf& { return (i & 1) ? sl.item_edge_odd : sl.item_edge; };
  auto set_pos = [&](size_t i) -> DevBuf& { return (i & 1) ? sl.item_pos_odd : sl.item_pos; };
  // Launches expected after the first window when the largest testnum is not
  // known yet (from the slot's earlier calls or the caller's bound): they only
  // size the buffer sets, so that the later enqueue does not have to grow
  // buffers that launches in flight still use.
  std::vector<Launch> expected;
  auto enqueue = [&](size_t from) -> int {
    int cap[2] = {0, 0};
    auto size_launch = [&](size_t i, const Launch& p) {
      cap[i & 1] = std::max(cap[i & 1], (int)frame_stride((long long)(p.e1 - p.e0) * (p.hi - p.lo)));
    };
    for (size_t i = from; i < plan.size(); ++i) size_launch(i, plan[i]);
    for (size_t j = 0; j < expected.size(); ++j) size_launch(plan.size() + j, expected[j]);
    size_t fbytes[2], ibytes[2];
    for (int k = 0; k < 2; ++k) {
      fbytes[k] = sizeof(float) * 12 * nd * (size_t)std::max(cap[k], 1);
      ibytes[k] = sizeof(int) * (size_t)std::max(cap[k], 1);
    }
    // per launch: [0] item count, [1] work counter
    const size_t cbytes = sizeof(int) * 2 * std::max<size_t>(plan.size() + expected.size(), 32);
    expected.clear();
    bool grow = cbytes > sl.counters.bytes;
    for (int k = 0; k < 2; ++k)
      grow = grow || (cap[k] > 0 && (fbytes[k] > set_frames(k).bytes || ibytes[k] > set_edge(k).bytes));
    if (from > 0 && grow) {
      // earlier launches still use the buffers to be replaced; their item
      // counts are kept on the host
      GPRM_HIP(hipStreamSynchronize(st), "sync");
      early.resize(2 * from);
      GPRM_HIP(hipMemcpy(early.data(), sl.counters.p, sizeof(int) * early.size(), hipMemcpyDeviceToHost), "copy");
    }
    for (int k = 0; k < 2; ++k) {
      if (cap[k] == 0) continue;
      GPRM_HIP(set_frames(k).ensure(fbytes[k]), "hipMalloc frames");
      GPRM_HIP(set_edge(k).ensure(ibytes[k]), "hipMalloc items");
      GPRM_HIP(set_pos(k).ensure(ibytes[k]), "hipMalloc items");
    }
    GPRM_HIP(sl.counters.ensure(cbytes), "hipMalloc counters");
    GPRM_HIP(sl.hist.ensure(sizeof(float) * 4 * kMaxIter * (size_t)c->shape.grid * c->shape.block),
             "hipMalloc history");
    if (plan.size() > 1 && !sl.prep) GPRM_HIP(hipStreamCreateWithFlags(&sl.prep, hipStreamNonBlocking), "hipStreamCreate");
    while (sl.ev_fk.size() < plan.size()) {
      hipEvent_t a, b;
      GPRM_HIP(hipEventCreateWithFlags(&a, hipEventDisableTiming), "hipEventCreate");
      sl.ev_fk.push_back(a);
      GPRM_HIP(hipEventCreateWithFlags(&b, hipEventDisableTiming), "hipEventCreate");
      sl.ev_sva.push_back(b);
    }
    if (!sl.ev_first) GPRM_HIP(hipEventCreateWithFlags(&sl.ev_first, hipEventDisableTiming), "hipEventCreate");
    bool prep_started = false;
    for (size_t i = from; i < plan.size(); ++i) {
      const Launch& p = plan[i];
      int* cnt = sl.counters.as<int>() + 2 * i;
      hipStream_t ps = st;
      if (i == 0) {
        GPRM_HIP(launch_fill_u32(sl.counters.as<unsigned>(), 0u, 2, st), "fill");
      } else {
        ps = sl.prep;
        if (!prep_started) {
          prep_started = true;
          GPRM_HIP(hipStreamWaitEvent(ps, sl.ev_first, 0), "stream wait");
          GPRM_HIP(launch_fill_u32(sl.counters.as<unsigned>() + 2 * i, 0u, 2 * (long long)(plan.size() - i), ps),
                   "fill");
        }
        if (i >= 2) GPRM_HIP(hipStreamWaitEvent(ps, sl.ev_sva[i - 2], 0), "stream wait");
      }
      int* items_e = set_edge(i).as<int>();
      int* items_p = set_pos(i).as<int>();
      float* frames = set_frames(i).as<float>();
      GPRM_HIP(launch_edge_compact(ea, p.e0, p.e1, p.lo, p.hi, items_e, items_p, cnt, ps),
               "edge compaction launch");
      // item capacity and frame plane stride
      const int pcap = (int)frame_stride((long long)(p.e1 - p.e0) * (p.hi - p.lo));
      size_t ev = 0;
      GPRM_HIP(prof_begin(c, sl, 1, ps, &ev), "event");
      GPRM_HIP(launch_edge_fk(c->rb, ea, items_e, items_p, cnt, pcap, frames, ps), "edge kinematics launch");
      GPRM_HIP(prof_end(c, sl, ev, ps), "event");
      if (i == 0) {
        GPRM_HIP(hipEventRecord(sl.ev_first, st), "event");
      } else {
        GPRM_HIP(hipEventRecord(sl.ev_fk[i], ps), "event");
        GPRM_HIP(hipStreamWaitEvent(st, sl.ev_fk[i], 0), "stream wait");
      }
      SvaArgs a{};
      a.img = c->img;
      a.ndof = nd;
      a.frames = frames;
      a.fstride = pcap;
      a.hist = sl.hist.as<float4>();
      a.nitems = -1;
      a.nitems_dev = cnt;
      a.counter = reinterpret_cast<unsigned*>(cnt + 1);
      a.item_edge = items_e;
      a.edge_res = d_res;
      a.pairs_per_unit = unit_pairs(c, pcap);
      a.units_per_item = (c->img.M + a.pairs_per_unit - 1) / a.pairs_per_unit;
      a.tally = tally_ptr(c, sl);
      a.static_first = 1;
      GPRM_HIP(prof_begin(c, sl, 0, st, &ev), "event");
      GPRM_HIP(launch_sva(a, 1, launch_shape(c, (long long)((pcap + 63) / 64) * a.units_per_item), st),
               "separating-vector launch");
      GPRM_HIP(prof_end(c, sl, ev, st), "event");
      GPRM_HIP(hipEventRecord(sl.ev_sva[i], st), "event");
    }
    return GPRM_OK;
  };
  int rc = GPRM_OK;
  int maxT = host_maxT;
  if (host_maxT >= 0) {
    // the whole plan is known: one enqueue sizes and launches it
    for (const auto& w : edge_windows(host_maxT, (long long)N * host_maxT * c->img.M)) add_window(w.first, w.second);
  } else {
    add_window(0, kFirstWindow);
    const int hint = std::max(sl.maxT_hint, maxT_hint);
    std::swap(plan, expected);
    for (const auto& w : edge_windows(hint))
      if (w.first >= kFirstWindow) add_window(w.first, w.second);
    std::swap(plan, expected);
  }
  if ((rc = enqueue(0)) != GPRM_OK) return rc;
  unsigned long long total = 0;
  if (need_max) {
    // issued after the first round so that the device gets that round's
    // launches first
    GPRM_HIP(hipStreamWaitEvent(sl.side, sl.ev_prologue, 0), "stream wait");
    GPRM_HIP(hipMemcpyAsync(sl.pinned, sl.misc.p, kMiscReadback, hipMemcpyDeviceToHost, sl.side), "copy max");
    GPRM_HIP(hipStreamSynchronize(sl.side), "sync");
    std::memcpy(&maxT, static_cast<char*>(sl.pinned) + kMiscMaxT, 4);
    std::memcpy(&total, static_cast<char*>(sl.pinned) + kMiscTotal, 8);
  }
  sl.maxT_hint = std::max(sl.maxT_hint, maxT);
  if (host_maxT < 0) {
    const size_t first = plan.size();
    for (const auto& w : edge_windows(maxT))
      if (w.first >= kFirstWindow) add_window(w.first, w.second);
    if ((rc = enqueue(first)) != GPRM_OK) return rc;
  }
  if (stats) {
    std::vector<int> h(2 * plan.size());
    if (!plan.empty()) {
      GPRM_HIP(hipMemcpyAsync(h.data(), sl.counters.p, sizeof(int) * h.size(), hipMemcpyDeviceToHost, st), "copy");
      GPRM_HIP(hipStreamSynchronize(st), "sync");
    }
    for (size_t i = 0; i < early.size(); ++i) h[i] = early[i];
    long long sched = 0;
    for (size_t i = 0; i < plan.size(); ++i) sched += h[2 * i];
    stats->samples_total = (long long)total;
    stats->samples_scheduled = sched;
    stats->max_testnum = maxT;
    stats->rounds = (int)plan.size();
  }
  return GPRM_OK;
}

// Host-side testnum (robotconfigspace.cpp:660-673), same float operations as
// the prologue kernel; gives the round plan without a device round trip.
long long host_testnum(const gprm_ctx* c, const float* qs, const float* qe, int N, int offset, int* maxT) {
  long long tot = 0;
  int mx = 0;
  for (int l = 0; l < N; ++l) {
    float norm = 0.0f;
    for (int i = 0; i < c->scene.ndof; ++i) {
      float diff = qe[l + (long long)offset * i] - qs[l + (long long)offset * i];
      norm += diff * diff;
    }
    norm = std::sqrt(norm);
    int T = (int)(norm / c->dq) + 1;
    tot += T;
    mx = std::max(mx, T);
  }
  *maxT = mx;
  return tot;
}

// Copy a caller SoA array (component i at [k + offset*i]) to device SoA with
// stride N, as the reference does (robotconfigspace.cpp:678-682).
int upload_soa(gprm_ctx* c, DevBuf& buf, const float* h, int N, int offset, hipStream_t st) {
  const int nd = c->scene.ndof;
  GPRM_HIP(buf.ensure(sizeof(float) * (size_t)nd * std::max(N, 1)), "hipMalloc");
  for (int i = 0; i < nd; ++i)
    GPRM_HIP(hipMemcpyAsync(buf.as<float>() + (size_t)N * i, h + (long long)offset * i, sizeof(float) * N,
                            hipMemcpyHostToDevice, st),
             "copy H2D");
  return GPRM_OK;
}

int check_ready(gprm_ctx* c) {
  if (!c) return GPRM_E_INVALID;
  if (!c->inited) return fail(c, GPRM_E_STATE, "context not initialised");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  return GPRM_OK;
}

}  // namespace

// Entry points for the planner (prm.cpp), which keeps one slot per worker
// and enqueues without synchronising.
int slot_acquire(gprm_ctx* c, int* idx) { return acquire_slot(c, idx); }
void slot_release(gprm_ctx* c, int idx) { c->slots[idx]->busy = false; }
int enqueue_points(gprm_ctx* c, Slot& sl, const float* d_q, long long offset, int n, int* d_res, hipStream_t st,
                   const unsigned char* skip) {
  return run_points(c, sl, d_q, offset, n, d_res, st, skip);
}
int enqueue_edges(gprm_ctx* c, Slot& sl, const float* d_qs, const float* d_qe, long long offset, int N, int* d_res,
                  hipStream_t st, int maxT_bound, const unsigned char* skip) {
  return run_edges(c, sl, d_qs, d_qe, offset, N, d_res, st, maxT_bound, nullptr, 0, skip);
}

// Grow the slot's buffers for point checks of n items and edge checks of N
// edges whose rounds are planned from maxT_bound (the sizes run_points and
// run_edges ensure), with headroom.  Work already enqueued may still use the
// buffers, so the device is synchronised before any of them is replaced.
int reserve_slot(gprm_ctx* c, Slot& sl, int n_points, int N, int maxT_bound) {
  const int nd = c->scene.ndof;
  struct Need { DevBuf* b; size_t bytes; };
  std::vector<Need> need;
  const size_t hist = sizeof(float) * 4 * kMaxIter * (size_t)c->shape.grid * c->shape.block;
  need.push_back({&sl.hist, hist});
  if (n_points > 0) {
    const int chunk = std::min(n_points, item_budget(c));
    need.push_back({&sl.frames, sizeof(float) * 12 * nd * (size_t)frame_stride(chunk)});
    need.push_back({&sl.status, sizeof(int) * (size_t)chunk});
    need.push_back({&sl.counters, sizeof(unsigned) * (size_t)((n_points + chunk - 1) / chunk)});
    need.push_back({&sl.groups, sizeof(int) * (size_t)((chunk + 63) / 64)});
  }
  if (N > 0) {
    // the launches of run_edges's plan: windows split into edge ranges,
    // buffer set i & 1 for launch i
    long long cap[2] = {0, 0};
    size_t launches = 0;
    for (const auto& w : edge_windows(maxT_bound, (long long)N * maxT_bound * c->img.M)) {
      const int per = std::max(1, item_budget(c) / (w.second - w.first));
      for (int e0 = 0; e0 < N; e0 += per, ++launches)
        cap[launches & 1] = std::max(cap[launches & 1],
                                     frame_stride((long long)(std::min(N, e0 + per) - e0) * (w.second - w.first)));
    }
    DevBuf* fr[2] = {&sl.frames, &sl.frames_odd};
    DevBuf* ie[2] = {&sl.item_edge, &sl.item_edge_odd};
    DevBuf* ip[2] = {&sl.item_pos, &sl.item_pos_odd};
    for (int k = 0; k < 2; ++k) {
      if (cap[k] == 0) continue;
      need.push_back({fr[k], sizeof(float) * 12 * nd * (size_t)cap[k]});
      need.push_back({ie[k], sizeof(int) * (size_t)cap[k]});
      need.push_back({ip[k], sizeof(int) * (size_t)cap[k]});
    }
    need.push_back({&sl.counters, sizeof(int) * 2 * std::max<size_t>(launches, 32)});
    need.push_back({&sl.testnum, sizeof(int) * (size_t)N});
    need.push_back({&sl.stride, sizeof(int) * (size_t)N});
    need.push_back({&sl.misc, 64});
  }
  bool grow = false;
  for (const auto& x : need) grow = grow || x.bytes > x.b->bytes;
  if (!grow) return GPRM_OK;
  GPRM_HIP(hipDeviceSynchronize(), "sync before growing buffers");
  for (const auto& x : need)
    if (x.bytes > x.b->bytes) GPRM_HIP(x.b->ensure(x.bytes + x.bytes / 2), "hipMalloc planner buffers");
  return GPRM_OK;
}
const unsigned long long* edge_sample_total(Slot& sl) {
  return reinterpret_cast<const unsigned long long*>(sl.misc.as<char>() + kMiscTotal);
}
const int* edge_max_testnum(Slot& sl) { return reinterpret_cast<const int*>(sl.misc.as<char>() + kMiscMaxT); }

}  // namespace gprm

using namespace gprm;

extern "C" {

int gprm_create_from_dir(const char* path, float dq, int nbuf, gprm_ctx** out) {
  if (!path || !out) return fail(nullptr, GPRM_E_INVALID, "null argument");
  *out = nullptr;
  Scene s;
  std::string err = load_dir(path, &s);
  if (!err.empty()) return fail(nullptr, GPRM_E_SCENE, err);
  return new_ctx(std::move(s), dq, nbuf, out);
}

int gprm_create(int ndof, const float* a, const float* alpha, const float* q, const float* d,
                const int* types, int N, const int* sys, const int* n, const int* m,
                const float* vertices, const int* dsp, const int* cnt, const int* dest,
                const int* from, const int* to, int M, const float* mins, const float* maxs,
                float dq, int nbuf, gprm_ctx** out) {
  if (!out || !a || !alpha || !q || !d || !types || !sys || !n || !m || !vertices || !dsp || !cnt ||
      !from || !to || !mins || !maxs)
    return fail(nullptr, GPRM_E_INVALID, "null argument");
  *out = nullptr;
  if (ndof < 1 || ndof > kMaxDof || N < 1 || M < 1) return fail(nullptr, GPRM_E_INVALID, "bad sizes");
  Scene s;
  s.ndof = ndof;
  for (int i = 0; i < ndof; ++i) {
    s.a[i] = a[i]; s.alpha[i] = alpha[i]; s.q[i] = q[i]; s.d[i] = d[i]; s.types[i] = types[i];
    s.mins[i] = mins[i]; s.maxs[i] = maxs[i];
  }
  long long vo = 0, eo = 0;
  for (int k = 0; k < N; ++k) {
    if (n[k] < 1 || m[k] < 0) return fail(nullptr, GPRM_E_INVALID, "bad polytope size");
    Polytope p;
    p.verts.assign(vertices + 3 * vo, vertices + 3 * (vo + n[k]));
    p.dsp.assign(dsp + vo, dsp + vo + n[k]);
    p.cnt.assign(cnt + vo, cnt + vo + n[k]);
    if (m[k] > 0 && !dest) return fail(nullptr, GPRM_E_INVALID, "null dest");
    if (m[k] > 0) p.dest.assign(dest + eo, dest + eo + m[k]);
    vo += n[k];
    eo += m[k];
    s.polys.push_back(std::move(p));
    s.sys.push_back(sys[k]);
  }
  s.from.assign(from, from + M);
  s.to.assign(to, to + M);
  return new_ctx(std::move(s), dq, nbuf, out);
}

int gprm_init(gprm_ctx* c, int rank, int size) {
  if (!c) return GPRM_E_INVALID;
  std::lock_guard<std::mutex> lk(c->mu);
  if (size < 1 || rank < 0 || rank >= size) return fail(c, GPRM_E_INVALID, "bad rank/size");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1) return fail(c, GPRM_E_HIP, "no HIP device available");
  // robotconfigspace.cpp:229: device = rank*devcount/size
  c->device = (int)((long long)rank * count / size);
  GPRM_HIP(hipSetDevice(c->device), "hipSetDevice");
  if (noalloc_setup(c->device) != GPRM_OK)
    return fail(c, GPRM_E_HIP, "hipGetSymbolAddress of the no-allocation scratch array failed");
  std::vector<char> image;
  int rc = build_image(c, &image);
  if (rc != GPRM_OK) return rc;
  GPRM_HIP(c->img_buf.ensure(image.size()), "hipMalloc scene");
  GPRM_HIP(hipMemcpy(c->img_buf.p, image.data(), image.size(), hipMemcpyHostToDevice), "copy scene");
  c->img.blob = c->img_buf.as<int4>();
  c->shape = sva_default_shape(c->img);
  GPRM_HIP(sva_shape_finish(c->img, &c->shape, c->device), "separating-vector launch shape");
  c->small_calls = point_small_prepare(c->img) == hipSuccess;
  c->inited = true;
  return GPRM_OK;
}

void gprm_destroy(gprm_ctx* c) {
  if (!c) return;
  if (c->device >= 0) (void)hipSetDevice(c->device);
  for (auto& s : c->slots)
    if (s->stream) (void)hipStreamSynchronize(s->stream);
  delete c;
}

int gprm_indicator(gprm_ctx* c, const float* q, int* res, int N, int offset) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (N < 0 || (N > 0 && (!q || !res)) || offset < 0) return fail(c, GPRM_E_INVALID, "bad arguments");
  if (N == 0) return GPRM_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  int si;
  if ((rc = acquire_slot(c, &si))) return rc;
  Slot& sl = *c->slots[si];
  struct Release { Slot& s; ~Release() { s.busy = false; } } rel{sl};
  hipStream_t st = sl.stream;
  if (N <= kSmallMax && c->small_calls && !c->profile) {
    // one launch with the configurations as arguments, codes written to
    // pinned memory: the reference's per-candidate indicator(q) calls
    // (prmsolver.cpp:2611) pay one round trip
    if (!sl.small_out)
      GPRM_HIP(hipHostMalloc((void**)&sl.small_out, sizeof(int) * kSmallMax, hipHostMallocDefault), "hipHostMalloc");
    SmallPoints sp{};
    sp.n = N;
    for (int k = 0; k < N; ++k)
      for (int i = 0; i < c->scene.ndof; ++i) sp.q[k][i] = q[k + (long long)offset * i];
    GPRM_HIP(launch_point_small(c->img, c->rb, sp, sl.small_out, st), "point-check launch");
    GPRM_HIP(hipStreamSynchronize(st), "sync");
    std::memcpy(res, sl.small_out, sizeof(int) * N);
    return GPRM_OK;
  }
  if ((rc = prof_prepare(c, sl, st))) return rc;
  if ((rc = upload_soa(c, sl.qs, q, N, offset, st))) return rc;
  GPRM_HIP(sl.res.ensure(sizeof(int) * (size_t)N), "hipMalloc res");
  if ((rc = run_points(c, sl, sl.qs.as<float>(), N, N, sl.res.as<int>(), st))) return rc;
  GPRM_HIP(hipMemcpyAsync(res, sl.res.p, sizeof(int) * N, hipMemcpyDeviceToHost, st), "copy D2H");
  GPRM_HIP(hipStreamSynchronize(st), "sync");
  return prof_harvest(c, sl);
}

int gprm_indicator1(gprm_ctx* c, const float* q, int* code) {
  if (!q || !code) return c ? fail(c, GPRM_E_INVALID, "null argument") : GPRM_E_INVALID;
  return gprm_indicator(c, q, code, 1, 1);
}

int gprm_indicator_dev(gprm_ctx* c, const float* d_q, int* d_res, int N, long long offset, void* stream) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (N < 0 || (N > 0 && (!d_q || !d_res)) || offset < 0) return fail(c, GPRM_E_INVALID, "bad arguments");
  if (N == 0) return GPRM_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  int si;
  if ((rc = acquire_slot(c, &si))) return rc;
  Slot& sl = *c->slots[si];
  struct Release { Slot& s; ~Release() { s.busy = false; } } rel{sl};
  hipStream_t st = stream ? (hipStream_t)stream : sl.stream;
  if ((rc = prof_prepare(c, sl, st))) return rc;
  if ((rc = run_points(c, sl, d_q, offset, N, d_res, st))) return rc;
  GPRM_HIP(hipStreamSynchronize(st), "sync");
  return prof_harvest(c, sl);
}

int gprm_kinematics_dev(gprm_ctx* c, const float* d_q, float* d_frames, int N, long long offset, void* stream) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (N < 0 || (N > 0 && (!d_q || !d_frames)) || offset < 0) return fail(c, GPRM_E_INVALID, "bad arguments");
  if (N == 0) return GPRM_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  int si;
  if ((rc = acquire_slot(c, &si))) return rc;
  Slot& sl = *c->slots[si];
  struct Release { Slot& s; ~Release() { s.busy = false; } } rel{sl};
  hipStream_t st = stream ? (hipStream_t)stream : sl.stream;
  const int nd = c->scene.ndof;
  const int chunk = std::min(N, item_budget(c));
  GPRM_HIP(sl.frames.ensure(sizeof(float) * 12 * nd * (size_t)frame_stride(chunk)), "hipMalloc frames");
  GPRM_HIP(sl.status.ensure(sizeof(int) * (size_t)chunk), "hipMalloc words");
  for (int c0 = 0; c0 < N; c0 += chunk) {
    const int cn = std::min(chunk, N - c0);
    // the kinematics kernel of the point checks, then its frames unpacked
    GPRM_HIP(launch_fk_points(c->rb, d_q + c0, offset, cn, sl.frames.as<float>(), frame_stride(cn),
                              sl.status.as<int>(), 0, nullptr, st),
             "kinematics launch");
    GPRM_HIP(launch_frames_unpack(c->rb, sl.frames.as<float>(), frame_stride(cn), cn,
                                  d_frames + (size_t)c0 * (nd + 1) * 12, st),
             "frames launch");
  }
  GPRM_HIP(hipStreamSynchronize(st), "sync");
  return GPRM_OK;
}

int gprm_indicator2_async(gprm_ctx* c, const float* qs, const float* qe, int* res, int N, int offset,
                          int* request) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (N < 0 || (N > 0 && (!qs || !qe || !res)) || offset < 0 || !request)
    return fail(c, GPRM_E_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  int si;
  if ((rc = acquire_slot(c, &si))) return rc;
  Slot& sl = *c->slots[si];
  hipStream_t st = sl.stream;
  int maxT = 0;
  long long total = host_testnum(c, qs, qe, N, offset, &maxT);
  if (N > 0) {
    rc = prof_prepare(c, sl, st);
    if (!rc) rc = upload_soa(c, sl.qs, qs, N, offset, st);
    if (!rc) rc = upload_soa(c, sl.qe, qe, N, offset, st);
    if (!rc && sl.res.ensure(sizeof(int) * (size_t)N) != hipSuccess) rc = fail(c, GPRM_E_HIP, "hipMalloc res");
    if (!rc) rc = run_edges(c, sl, sl.qs.as<float>(), sl.qe.as<float>(), N, N, sl.res.as<int>(), st, maxT, nullptr);
    if (rc) {
      sl.busy = false;
      return rc;
    }
  }
  sl.host_res = res;
  sl.N = N;
  int id = ++c->next_request;
  c->requests[id] = si;
  *request = id;
  c->numthreads_all += total;  // robotconfigspace.cpp:720
  return GPRM_OK;
}

int gprm_indicator2_wait(gprm_ctx* c, int request) {
  int rc = check_ready(c);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->requests.find(request);
  if (it == c->requests.end()) return fail(c, GPRM_E_REQUEST, "unknown request id");
  Slot& sl = *c->slots[it->second];
  c->requests.erase(it);
  struct Release { Slot& s; ~Release() { s.busy = false; } } rel{sl};
  if (sl.N > 0)
    GPRM_HIP(hipMemcpyAsync(sl.host_res, sl.res.p, sizeof(int) * sl.N, hipMemcpyDeviceToHost, sl.stream),
             "copy D2H");
  GPRM_HIP(hipStreamSynchronize(sl.stream), "sync");
  return prof_harvest(c, sl);
}

int gprm_indicator2(gprm_ctx* c, const float* qs, const float* qe, int* res, int N, int offset) {
  int req = 0;
  int rc = gprm_indicator2_async(c, qs, qe, res, N, offset, &req);
  if (rc) return rc;
  return gprm_indicator2_wait(c, req);
}

int gprm_indicator2_dev(gprm_ctx* c, const float* d_qs, const float* d_qe, int* d_res, int N,
                        long long offset, void* stream, gprm_edge_stats* stats) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (N < 0 || (N > 0 && (!d_qs || !d_qe || !d_res)) || offset < 0) return fail(c, GPRM_E_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  int si;
  if ((rc = acquire_slot(c, &si))) return rc;
  Slot& sl = *c->slots[si];
  struct Release { Slot& s; ~Release() { s.busy = false; } } rel{sl};
  hipStream_t st = stream ? (hipStream_t)stream : sl.stream;
  gprm_edge_stats local;
  if ((rc = prof_prepare(c, sl, st))) return rc;
  if ((rc = run_edges(c, sl, d_qs, d_qe, offset, N, d_res, st, -1, &local))) return rc;
  GPRM_HIP(hipStreamSynchronize(st), "sync");
  c->numthreads_all += local.samples_total;
  if (stats) *stats = local;
  return prof_harvest(c, sl);
}

int gprm_indicator2_compressed(gprm_ctx* c, const float* qs, int M, const float* qe, int* res,
                               const int* posqe, const int* numqe, int offset) {
  if (!c) return GPRM_E_INVALID;
  if (M < 0 || (M > 0 && (!qs || !qe || !res || !posqe || !numqe)) || offset < 0)
    return fail(c, GPRM_E_INVALID, "bad arguments");
  const int nd = c->scene.ndof;
  long long E = 0;
  for (int i = 0; i < M; ++i) {
    if (numqe[i] < 0 || posqe[i] < 0) return fail(c, GPRM_E_INVALID, "bad posqe/numqe");
    E += numqe[i];
  }
  if (E > (1LL << 30)) return fail(c, GPRM_E_INVALID, "too many edges");
  std::vector<float> S((size_t)nd * E), T((size_t)nd * E);
  std::vector<int> idx(E), r(E);
  long long e = 0;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < numqe[i]; ++j, ++e) {
      idx[e] = posqe[i] + j;
      for (int d = 0; d < nd; ++d) {
        S[(size_t)E * d + e] = qs[i + (long long)offset * d];
        T[(size_t)E * d + e] = qe[idx[e] + (long long)offset * d];
      }
    }
  int rc = gprm_indicator2(c, S.data(), T.data(), r.data(), (int)E, (int)E);
  if (rc) return rc;
  for (long long k = 0; k < E; ++k) res[idx[k]] = r[k];
  return GPRM_OK;
}

int gprm_check_boundaries(gprm_ctx* c, const float* q, int* res, int N, int offset) {
  if (!c) return GPRM_E_INVALID;
  if (N < 0 || (N > 0 && (!q || !res)) || offset < 0) return fail(c, GPRM_E_INVALID, "bad arguments");
  for (int ix = 0; ix < N; ++ix) {
    int r = 0;
    for (int i = 0; i < c->scene.ndof; ++i) {
      float qi = q[ix + (long long)i * offset];
      if (qi < c->scene.mins[i] || qi > c->scene.maxs[i]) { r = 1; break; }
    }
    res[ix] = r;
  }
  return GPRM_OK;
}

int gprm_check_boundaries1(gprm_ctx* c, const float* q, int* out) {
  if (!c || !q || !out) return GPRM_E_INVALID;
  return gprm_check_boundaries(c, q, out, 1, 1);
}

float gprm_min(gprm_ctx* c, int i) { return (c && i >= 0 && i < c->scene.ndof) ? c->scene.mins[i] : NAN; }
float gprm_max(gprm_ctx* c, int i) { return (c && i >= 0 && i < c->scene.ndof) ? c->scene.maxs[i] : NAN; }
float gprm_deltaq(gprm_ctx* c) { return c ? c->dq : NAN; }
int gprm_dim(gprm_ctx* c) { return c ? c->scene.ndof : GPRM_E_INVALID; }
long long gprm_numthreads_all(gprm_ctx* c) { return c ? c->numthreads_all : GPRM_E_INVALID; }

int gprm_profile_enable(gprm_ctx* c, int enable) {
  if (!c) return GPRM_E_INVALID;
  std::lock_guard<std::mutex> lk(c->mu);
  c->profile = enable != 0;
  c->prof = gprm_profile_stats{};
  for (auto& s : c->slots) s->tally_pending = false;  // the next profiled call clears them
  return GPRM_OK;
}

int gprm_profile_read(gprm_ctx* c, gprm_profile_stats* out) {
  if (!c || !out) return GPRM_E_INVALID;
  std::lock_guard<std::mutex> lk(c->mu);
  for (auto& s : c->slots) {
    if (s->busy) continue;  // an asynchronous request in flight: folded by a later read
    int rc = prof_fold_tally(c, *s);
    if (rc) return rc;
  }
  *out = c->prof;
  return GPRM_OK;
}

int gprm_get_launch_shape(gprm_ctx* c, gprm_launch_shape* out) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!out) return fail(c, GPRM_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  *out = gprm_launch_shape{c->shape.block, c->shape.waves, c->shape.ppu, c->shape.grid};
  return GPRM_OK;
}

int gprm_set_launch_shape(gprm_ctx* c, const gprm_launch_shape* in) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!in) return fail(c, GPRM_E_INVALID, "null argument");
  SvaShape sh;
  sh.block = in->block;
  sh.waves = in->waves;
  sh.ppu = in->pairs_per_unit;
  hipError_t e = sva_shape_finish(c->img, &sh, c->device);
  if (e == hipErrorInvalidValue || e == hipErrorInvalidConfiguration)
    return fail(c, GPRM_E_INVALID, "launch shape not available for this scene");
  if (e != hipSuccess) return hip_fail(c, e, "launch shape");
  std::lock_guard<std::mutex> lk(c->mu);
  c->shape = sh;
  return GPRM_OK;
}

int gprm_autotune(gprm_ctx* c, const float* d_q, const float* d_qe, int n, long long offset, int reps,
                  gprm_launch_shape* best, float* ms_best) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!d_q || n < 1 || offset < n || reps < 1 || !best) return fail(c, GPRM_E_INVALID, "bad arguments");
  int* d_res = nullptr;
  GPRM_HIP(hipMalloc(&d_res, sizeof(int) * (size_t)n), "hipMalloc results");
  struct Free { int* p; ~Free() { (void)hipFree(p); } } fr{d_res};
  const gprm_launch_shape variants[] = {{256, 5, 0, 0}, {256, 4, 0, 0}, {1024, 4, 0, 0}};
  const int ppus[] = {32, 64, 128, 256};
  gprm_launch_shape keep{};
  if ((rc = gprm_get_launch_shape(c, &keep)) != GPRM_OK) return rc;
  double best_ms = 1e300;
  gprm_launch_shape found = keep;
  for (const auto& v : variants)
    for (int ppu : ppus) {
      gprm_launch_shape sh = v;
      sh.pairs_per_unit = ppu;
      if (gprm_set_launch_shape(c, &sh) != GPRM_OK) continue;  // not available for this scene
      // the median of reps timed runs after one warm-up
      std::vector<double> ts;
      for (int r = 0; r <= reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        rc = d_qe ? gprm_indicator2_dev(c, d_q, d_qe, d_res, n, offset, nullptr, nullptr)
                  : gprm_indicator_dev(c, d_q, d_res, n, offset, nullptr);
        const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc != GPRM_OK) {
          (void)gprm_set_launch_shape(c, &keep);
          return rc;
        }
        if (r > 0) ts.push_back(t);
      }
      std::sort(ts.begin(), ts.end());
      const double ms = ts[ts.size() / 2];
      if (ms < best_ms) {
        best_ms = ms;
        (void)gprm_get_launch_shape(c, &found);
      }
    }
  if ((rc = gprm_set_launch_shape(c, &found)) != GPRM_OK) return rc;
  *best = found;
  if (ms_best) *ms_best = (float)best_ms;
  return GPRM_OK;
}

const char* gprm_last_error(gprm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int gprm_scene_generate(int workload, unsigned long long seed, const char* out_dir) {
  if (!out_dir) return fail(nullptr, GPRM_E_INVALID, "null argument");
  Scene s;
  std::string err = generate(workload, seed, &s);
  if (!err.empty()) return fail(nullptr, GPRM_E_INVALID, err);
  err = write_dir(s, out_dir);
  if (!err.empty()) return fail(nullptr, GPRM_E_SCENE, err);
  return GPRM_OK;
}

}  // extern "C"
