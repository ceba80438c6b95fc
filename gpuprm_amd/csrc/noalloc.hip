// noalloc.hip -- point checks into a fixed device scratch array, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// gprm_indicator_dev_noalloc runs the same kernels as gprm_indicator_dev
// (kinematics, persistent separating-vector kernel, finalize), but every
// buffer beyond the caller's inputs and outputs lives in `scratch`, one
// file-scope device array per GPU.  Its device address is looked up once per
// device, when a context is initialised there (gprm_init calls
// noalloc_setup with the device current); host code passes only that
// address to kernels.  The call only enqueues kernels on the caller's
// stream (no allocation, no memset node, no synchronisation), so it may also
// be captured into a HIP graph.  All calls on one GPU share the array: calls
// whose work overlaps in time on the same GPU (on different streams, or from
// several host threads through different contexts) race and must be
// serialised by the caller; calls in order on one stream do not.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gprm.h"
#include "context.h"
#include "runtime.h"

namespace gprm {
namespace {

// Layout of the scratch array (byte offsets, 256-aligned):
//   [0, 256)                        work counter of the separating-vector launch
//   [256, 256 + 4 kMaxItems)        result words, one per configuration
//   frames: kFramesBytes            link frames, 48 bytes per link and configuration
//   history: kHistBytes             per-lane iteration history, kMaxIter float4 per lane
constexpr int kMaxItems = 1 << 16;
constexpr size_t kCounterOff = 0;
constexpr size_t kWordsOff = 256;
constexpr size_t kFramesOff = kWordsOff + sizeof(int) * (size_t)kMaxItems;
constexpr size_t kFramesBytes = 24u << 20;  // 65536 configurations of up to 8 joints
constexpr size_t kHistOff = kFramesOff + kFramesBytes;
constexpr size_t kHistLanes = 327680;       // 256 CUs x 5 blocks x 256 threads
constexpr size_t kHistBytes = sizeof(float4) * kMaxIter * kHistLanes;
constexpr size_t kScratchBytes = kHistOff + kHistBytes;  // 74.3 MiB
static_assert(kFramesOff % 256 == 0 && kHistOff % 256 == 0, "scratch regions must stay 256-aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};  // device address of `scratch` per device

}  // namespace

// Look up the device address of the scratch array on `device` (current);
// once per device.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return GPRM_E_INVALID;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return GPRM_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return GPRM_E_HIP;
  g_ptr[device] = p;
  return GPRM_OK;
}

// Largest configuration count of one call for an arm of ndof joints.
int noalloc_max_items(int ndof) {
  const size_t per_item = sizeof(float) * 12 * (size_t)(ndof > 0 ? ndof : 1);
  size_t n = kFramesBytes / per_item / 64 * 64;
  return (int)(n < (size_t)kMaxItems ? n : (size_t)kMaxItems);
}

}  // namespace gprm

using namespace gprm;

extern "C" int gprm_indicator_dev_noalloc(gprm_ctx* c, const float* d_q, int* d_res, int N, long long offset,
                                          void* stream) {
  if (!c) return GPRM_E_INVALID;
  if (!c->inited) {
    c->err = "context not initialised";
    return GPRM_E_STATE;
  }
  if (N < 0 || (N > 0 && (!d_q || !d_res)) || offset < 0) {
    c->err = "bad arguments";
    return GPRM_E_INVALID;
  }
  const int nd = c->scene.ndof;
  const int cap = noalloc_max_items(nd);
  if (N > cap) {
    c->err = "gprm_indicator_dev_noalloc: " + std::to_string(N) + " configurations exceed its maximum of " +
             std::to_string(cap) + " for " + std::to_string(nd) + " joints";
    return GPRM_E_CAPACITY;
  }
  if (N == 0) return GPRM_OK;
  if (hipSetDevice(c->device) != hipSuccess) {
    c->err = "hipSetDevice failed";
    return GPRM_E_HIP;
  }
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->device >= 0 && c->device < kMaxDevices) base = g_ptr[c->device];
  }
  if (!base) {
    c->err = "scratch address not set up on this device";
    return GPRM_E_STATE;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = static_cast<hipStream_t>(stream);
  char* b = static_cast<char*>(base);
  unsigned* counter = reinterpret_cast<unsigned*>(b + kCounterOff);
  int* words = reinterpret_cast<int*>(b + kWordsOff);
  float* frames = reinterpret_cast<float*>(b + kFramesOff);
  float4* hist = reinterpret_cast<float4*>(b + kHistOff);
  const long long fstride = frame_stride(N);
  // the kinematics launch writes every word and zeroes the work counter
  hipError_t e = launch_fk_points(c->rb, d_q, offset, N, frames, fstride, words, 1, counter, st);
  SvaArgs a{};
  a.img = c->img;
  a.ndof = nd;
  a.frames = frames;
  a.fstride = fstride;
  a.hist = hist;
  a.nitems = N;
  a.counter = counter;
  a.pairs_per_unit = unit_pairs(c, N);
  a.units_per_item = (c->img.M + a.pairs_per_unit - 1) / a.pairs_per_unit;
  a.word = words;
  a.tally = nullptr;
  a.static_first = 1;
  SvaShape sh = launch_shape(c, (long long)((N + 63) / 64) * a.units_per_item);
  // the history holds kHistLanes lanes: a larger resident grid is cut
  if ((size_t)sh.grid * sh.block > kHistLanes) sh.grid = (int)(kHistLanes / sh.block);
  if (e == hipSuccess) e = launch_sva(a, 0, sh, st);
  if (e == hipSuccess) e = launch_finalize_points(words, d_res, N, st);
  if (e != hipSuccess) {
    c->err = std::string("gprm_indicator_dev_noalloc: ") + hipGetErrorString(e);
    return GPRM_E_HIP;
  }
  return GPRM_OK;
}
