// context.h -- the state behind a gprm_ctx (one scene on one device), shared
// by the translation units of the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gprm.h"
#include "runtime.h"
#include "scene.h"

namespace gprm {

// Order-position windows of the edge rounds for a largest testnum maxT:
// [0, 12), then widths growing fourfold, or one window [0, maxT) when
// pair_tests (a bound on the call's pair tests, -1 if unknown) is small
// (defined in gprm_api.cpp).
std::vector<std::pair<int, int>> edge_windows(int maxT, long long pair_tests = -1);

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t ensure(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Layout of a slot's `misc` words written by the edge prologue and read back
// into `pinned`: the largest testnum (int) and the sample total (u64).
constexpr size_t kMiscMaxT = 0;
constexpr size_t kMiscTotal = 16;
constexpr size_t kMiscReadback = 24;

// Buffers of one in-flight request (the reference keeps one buffer set and
// one stream per outstanding request, robotconfigspace.cpp:627-650).
struct Slot {
  hipStream_t stream = nullptr;
  // edge calls read the prologue's largest testnum back on a side stream
  // (into pinned memory) while the first round runs on the call's stream
  hipStream_t side = nullptr;
  hipEvent_t ev_prologue = nullptr;
  void* pinned = nullptr;  // 32 bytes: copy of misc[0, kMiscReadback)
  // Edge launches after the first compact their items and compute their
  // frames on the prep stream, into the buffer set of the launch's parity,
  // while the previous launch's separating-vector kernel runs: ev_first
  // follows the first launch's kinematics, ev_fk[i] launch i's kinematics,
  // ev_sva[i] its separating-vector kernel.
  hipStream_t prep = nullptr;
  hipEvent_t ev_first = nullptr;
  std::vector<hipEvent_t> ev_fk, ev_sva;
  DevBuf frames, status, item_edge, item_pos, counters, qs, qe, res, testnum, stride, misc, tally, hist;
  DevBuf groups;  // point calls: completed units per group of 64 items
  DevBuf frames_odd, item_edge_odd, item_pos_odd;  // buffer set of the odd edge launches
  int maxT_hint = 0;  // largest testnum of the slot's earlier edge calls (sizes buffers ahead)
  int* small_out = nullptr;  // pinned codes of small point-check calls (kSmallMax ints)
  int* host_res = nullptr;
  int N = 0;
  bool busy = false;
  // profiling: (start, end) event pairs of this call's launches; kind 0 = SVA, 1 = kinematics
  std::vector<hipEvent_t> ev_start, ev_end;
  std::vector<int> ev_kind;
  size_t ev_used = 0;
  bool tally_pending = false;
  ~Slot() {
    for (auto e : ev_start) (void)hipEventDestroy(e);
    for (auto e : ev_end) (void)hipEventDestroy(e);
    if (side) (void)hipStreamSynchronize(side);
    if (prep) (void)hipStreamSynchronize(prep);
    for (auto e : ev_fk) (void)hipEventDestroy(e);
    for (auto e : ev_sva) (void)hipEventDestroy(e);
    if (ev_first) (void)hipEventDestroy(ev_first);
    if (prep) (void)hipStreamDestroy(prep);
    if (ev_prologue) (void)hipEventDestroy(ev_prologue);
    if (pinned) (void)hipHostFree(pinned);
    if (small_out) {
      if (stream) (void)hipStreamSynchronize(stream);
      (void)hipHostFree(small_out);
    }
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace gprm

struct gprm_ctx;
namespace gprm {
struct PlanCache;  // the planner's device state between solves (prm.cpp)
// Pairs per work unit of a separating-vector launch over about n items, and
// the context's launch shape cut to what a launch of `units` work units can
// use (gprm_api.cpp).
int unit_pairs(const gprm_ctx* c, long long n);
SvaShape launch_shape(const gprm_ctx* c, long long units);
// For the planner (gprm_api.cpp): a slot held across calls, and the point
// and edge checks enqueued on `st` without synchronising.  enqueue_edges
// plans its rounds from maxT_bound (at least the largest testnum of the
// edges); skip (may be null) leaves edges unchecked.  edge_sample_total:
// device address of the last edge call's sum of testnum on that slot.
int slot_acquire(gprm_ctx* c, int* idx);
void slot_release(gprm_ctx* c, int idx);
int enqueue_points(gprm_ctx* c, Slot& sl, const float* d_q, long long offset, int n, int* d_res, hipStream_t st,
                   const unsigned char* skip);
int enqueue_edges(gprm_ctx* c, Slot& sl, const float* d_qs, const float* d_qe, long long offset, int N, int* d_res,
                  hipStream_t st, int maxT_bound, const unsigned char* skip);
const unsigned long long* edge_sample_total(Slot& sl);
const int* edge_max_testnum(Slot& sl);  // the same call's largest testnum
// Grow the slot's buffers (with headroom, after a device synchronisation)
// for point checks of n_points items and edge checks of N edges planned
// from maxT_bound, so that the enqueue calls above allocate nothing.
int reserve_slot(gprm_ctx* c, Slot& sl, int n_points, int N, int maxT_bound);
// Device address of the scratch array of gprm_indicator_dev_noalloc on
// `device` (the current device), looked up once per device (noalloc.hip).
int noalloc_setup(int device);
}  // namespace gprm

struct gprm_ctx {
  gprm::Scene scene;
  gprm::DevRobot rb{};
  float dq = 0.01f;
  int nbuf = 0;
  int device = -1;
  bool inited = false;
  bool small_calls = false;  // the scene fits the small point-check kernel's LDS
  gprm::DevBuf img_buf;
  gprm::SceneImg img{};
  gprm::SvaShape shape;
  std::vector<std::unique_ptr<gprm::Slot>> slots;
  std::map<int, int> requests;  // request id -> slot index
  int next_request = 0;
  long long numthreads_all = 0;
  bool profile = false;
  gprm_profile_stats prof{};
  std::string err;
  std::mutex mu;
  std::shared_ptr<gprm::PlanCache> plan_cache;
};
