// kernels.hip -- CDNA4 (gfx950) kernels of the exact collision-check path.
//
//   fk_points / fk_edges : forward kinematics, one lane per configuration,
//                          frames kept in VGPRs and written once to HBM.
//   sva                  : persistent separating-vector kernel; the scene is
//                          staged in LDS, every lane is an independent worker
//                          that steps one iteration of the Chung-Wang test per
//                          trip and refills itself from a wave-level work queue
//                          (ballot + one atomic per 64 items).
//   edge_prologue/compact: edge sample counts, sample order, and per-round
//                          compaction of the samples of still-open edges.
//
// Arithmetic follows the reference's host path operation by operation:
// the file is compiled with -ffp-contract=off (no FMA contraction), IEEE
// division and sqrt, denormals kept.  Comparisons keep the reference's
// direction so NaN behaves as it does there.
#include <hip/hip_runtime.h>

#include <cmath>

#include "glibc_trig.h"
#include "runtime.h"

namespace gprm {

namespace {

// ----------------------------------------------------------------------------
// geometry (prm/src/lib/collision4/geo4.h)
// ----------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
struct V2 {
  float x, y;
};
// trafo4 (geo4.h:22-24): columns 0..3, rows x,y,z.
struct T12 {
  float m[12];  // col c, row r at m[3*c + r]
};

__device__ __forceinline__ V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }

// sprod (geo4.h:130-132)
__device__ __forceinline__ float dot3(const V3& u, const V3& v) {
  return (u.x * v.x + u.y * v.y) + u.z * v.z;
}
// normalize(float4&) (geo4.h:165-171)
__device__ __forceinline__ float normalize3(V3& u) {
  float norm = __builtin_sqrtf((u.x * u.x + u.y * u.y) + u.z * u.z);
  float factor = 1.0f / norm;
  if (factor == __builtin_inff()) factor = 0.0f;
  u.x *= factor;
  u.y *= factor;
  u.z *= factor;
  return norm;
}
// normalize(float2&) (geo4.h:80-83), no zero guard
__device__ __forceinline__ void normalize2(V2& u) {
  float factor = 1.0f / __builtin_sqrtf(u.x * u.x + u.y * u.y);
  u.x *= factor;
  u.y *= factor;
}
// cross2 (geo4.h:66-68)
__device__ __forceinline__ float cross2(const V2& v, const V2& w) { return v.x * w.y - v.y * w.x; }

// trafo4::apply(u, Tu) (geo4.h:215-221)
__device__ __forceinline__ V3 t_apply(const T12& t, const V3& u) {
  const float* c = t.m;
  V3 r;
  r.x = ((c[0] * u.x + c[3] * u.y) + c[6] * u.z) + c[9];
  r.y = ((c[1] * u.x + c[4] * u.y) + c[7] * u.z) + c[10];
  r.z = ((c[2] * u.x + c[5] * u.y) + c[8] * u.z) + c[11];
  return r;
}
// trafo4::apply_rot(u) (geo4.h:223-230)
__device__ __forceinline__ V3 t_apply_rot(const T12& t, const V3& u) {
  const float* c = t.m;
  V3 r;
  r.x = (c[0] * u.x + c[3] * u.y) + c[6] * u.z;
  r.y = (c[1] * u.x + c[4] * u.y) + c[7] * u.z;
  r.z = (c[2] * u.x + c[5] * u.y) + c[8] * u.z;
  return r;
}
// trafo4::apply_rot_inv(u, Rtu) (geo4.h:249-255)
__device__ __forceinline__ V3 t_apply_rot_inv(const T12& t, const V3& u) {
  const float* c = t.m;
  V3 r;
  r.x = (c[0] * u.x + c[1] * u.y) + c[2] * u.z;
  r.y = (c[3] * u.x + c[4] * u.y) + c[5] * u.z;
  r.z = (c[6] * u.x + c[7] * u.y) + c[8] * u.z;
  return r;
}
// trafo4::set(a, alpha, q, d) (geo4.h:194-203), with ca = cos(alpha) and
// sa = sin(alpha) given (the twist is constant per joint; DevRobot holds them)
__device__ __forceinline__ T12 t_set_cs(float a, float ca, float sa, float q, float d) {
  float cq = gprm_trig::cosf_ref(q), sq = gprm_trig::sinf_ref(q);
  T12 t;
  t.m[0] = cq;      t.m[3] = -sq;     t.m[6] = 0.0f;  t.m[9] = a;
  t.m[1] = sq * ca; t.m[4] = cq * ca; t.m[7] = -sa;   t.m[10] = -sa * d;
  t.m[2] = sq * sa; t.m[5] = cq * sa; t.m[8] = ca;    t.m[11] = ca * d;
  return t;
}
// this = T * this (trafo4::lapply, geo4.h:272-279)
__device__ __forceinline__ T12 t_lapply(const T12& self, const T12& T) {
  T12 r;
  for (int c = 0; c < 3; ++c) {
    V3 col = t_apply_rot(T, V3{self.m[3 * c], self.m[3 * c + 1], self.m[3 * c + 2]});
    r.m[3 * c] = col.x; r.m[3 * c + 1] = col.y; r.m[3 * c + 2] = col.z;
  }
  V3 col = t_apply(T, V3{self.m[9], self.m[10], self.m[11]});
  r.m[9] = col.x; r.m[10] = col.y; r.m[11] = col.z;
  return r;
}

// Kinematics<ndof>::calculate (kinematics.h:27-37); frames 1..ndof of item
// to the frame array, link-major and quad-planar: float4 j (0..2) of link
// frame i of item at out4[(3 i + j) stride + item], so each quad of one link
// is a contiguous array over the items.
// T0 = set(0, 0, 0, 0) (kinematics.h:27-37; trafo4::set, geo4.h:32-47),
// from cos/sin(0) held in rb
__device__ __forceinline__ T12 t_base(const DevRobot& rb) {
  const float cq = rb.c0, sq = rb.s0, ca = rb.c0, sa = rb.s0, a = 0.0f, d = 0.0f;
  T12 T;
  T.m[0] = cq;      T.m[3] = -sq;     T.m[6] = 0.0f;  T.m[9] = a;
  T.m[1] = sq * ca; T.m[4] = cq * ca; T.m[7] = -sa;   T.m[10] = -sa * d;
  T.m[2] = sq * sa; T.m[5] = cq * sa; T.m[8] = ca;    T.m[11] = ca * d;
  return T;
}

template <int NDOF>
__device__ __forceinline__ void fk_store(const DevRobot& rb, const float* q, float* __restrict__ out, int item,
                                         long long stride) {
  const int nd = NDOF > 0 ? NDOF : rb.ndof;
  T12 T = t_base(rb);
#pragma unroll
  for (int i = 0; i < (NDOF > 0 ? NDOF : kMaxDofDev); ++i) {
    if (NDOF == 0 && i >= nd) break;
    float qset = rb.q[i], dset = rb.d[i];
    if (rb.types[i] == 0) qset = q[i];
    else dset = q[i];
    T12 A = t_set_cs(rb.a[i], rb.ca[i], rb.sa[i], qset, dset);
    T = t_lapply(A, T);
    // quad-planar: the lanes of a wave store each quad of 64 consecutive
    // items as one contiguous 1 KB span
    float4* o = reinterpret_cast<float4*>(out) + (size_t)i * 3 * stride + item;
    o[0] = make_float4(T.m[0], T.m[1], T.m[2], T.m[3]);
    o[stride] = make_float4(T.m[4], T.m[5], T.m[6], T.m[7]);
    o[2 * stride] = make_float4(T.m[8], T.m[9], T.m[10], T.m[11]);
  }
}

// Frame of link sys (1..ndof) of item from the frame array: link-major and
// quad-planar, float4 j (0..2) of link s of item i at
// frames4[((s - 1) 3 + j) stride + i].
__device__ __forceinline__ T12 load_frame(const float* __restrict__ frames, int item, long long stride, int sys) {
  const float4* f = reinterpret_cast<const float4*>(frames) + (size_t)(sys - 1) * 3 * stride + item;
  const float4 a = f[0], b = f[stride], c = f[2 * stride];
  T12 t;
  t.m[0] = a.x; t.m[1] = a.y; t.m[2] = a.z;  t.m[3] = a.w;
  t.m[4] = b.x; t.m[5] = b.y; t.m[6] = b.z;  t.m[7] = b.w;
  t.m[8] = c.x; t.m[9] = c.y; t.m[10] = c.z; t.m[11] = c.w;
  return t;
}

// `norm < collision_eps` and `dp >= -collision_eps` compare a float with the
// double 1e-6 (collision4.h:20,242,296).  float(1e-6) = 0x1.0c6f7ap-20 lies
// below 1e-6, so both become exact float comparisons with it.
constexpr float kEpsBelow = 0x1.0c6f7ap-20f;

// ----------------------------------------------------------------------------
// kinematics kernels
// ----------------------------------------------------------------------------
// Point mode result word of one configuration: ((pair+1) << 4) | (code+1)
// of the first colliding pair in CRS order (updated by atomicMin), kWordFree
// while none is known; out of bounds is (0 << 4) | (1+1).
constexpr int kWordFree = 0x7fffffff;
constexpr int kWordOutOfBounds = 2;

template <int NDOF>
__global__ void __launch_bounds__(kFkBlock)
    fk_points_kernel(DevRobot rb, const float* __restrict__ q, long long offset, int n,
                     float* __restrict__ frames, long long fstride, int* __restrict__ word, int check_bounds,
                     unsigned* __restrict__ zero_counter, const unsigned char* __restrict__ skip,
                     int* __restrict__ group_done) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  // the separating-vector launch that follows claims units from this counter
  // and counts completed units per group of 64 items in group_done
  if (i == 0 && zero_counter != nullptr) *zero_counter = 0u;
  if (group_done != nullptr && (i & 63) == 0 && i < n) group_done[i >> 6] = 0;
  if (i >= n) return;
  const int nd = NDOF > 0 ? NDOF : rb.ndof;
  float qq[kMaxDofDev];
#pragma unroll
  for (int j = 0; j < (NDOF > 0 ? NDOF : kMaxDofDev); ++j)
    if (NDOF > 0 || j < nd) qq[j] = q[i + offset * j];
  if (skip != nullptr && skip[i]) {
    word[i] = kWordOutOfBounds;
  } else if (check_bounds) {
    // check_boundaries(q) (robotconfigspace.cpp:795-802)
    int st = 0;
    for (int j = 0; j < nd; ++j)
      if (qq[j] < rb.mins[j] || qq[j] > rb.maxs[j]) { st = 1; break; }
    word[i] = st ? kWordOutOfBounds : kWordFree;
  } else {
    word[i] = kWordFree;
  }
  fk_store<NDOF>(rb, qq, frames, i, fstride);
}

// The frames of n configurations as Kinematics<ndof>::calculate leaves its
// trafos (kinematics.h:28-37): the base frame and the link frames the
// kinematics kernel stored, unpacked to out[(i (ndof + 1) + s) 12 + m] with
// m = 3 column + row (trafo4's four columns of x, y, z).
__global__ void __launch_bounds__(256)
    frames_unpack_kernel(DevRobot rb, const float* __restrict__ frames, long long fstride, int n,
                         float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* o = out + (size_t)i * (rb.ndof + 1) * 12;
  const T12 T0 = t_base(rb);
  for (int m = 0; m < 12; ++m) o[m] = T0.m[m];
  for (int s = 1; s <= rb.ndof; ++s) {
    const T12 t = load_frame(frames, i, fstride, s);
    for (int m = 0; m < 12; ++m) o[12 * s + m] = t.m[m];
  }
}

// codes[i] from the point mode result word.
__global__ void finalize_points_kernel(const int* __restrict__ word, int* __restrict__ codes, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int w = word[i];
  codes[i] = w == kWordFree ? 0 : (w & 15) - 1;
}

// Frames of item i of an edge round: sample (pos * stride) mod T of its
// edge, interpolated as robotconfigspace.cpp:406-420 does (c1 = s/(T-1),
// c2 = 1 - c1, q = c1 qs + c2 qe), then the kinematics.
template <int NDOF>
__device__ __forceinline__ void fk_edge_item(const DevRobot& rb, const EdgeArgs& ea, const int* __restrict__ item_edge,
                                             const int* __restrict__ item_pos, int i, float* __restrict__ frames,
                                             long long fstride) {
  const int nd = NDOF > 0 ? NDOF : rb.ndof;
  const int e = item_edge[i];
  const int pos = item_pos[i];
  const int T = ea.testnum[e];
  const int s = (int)(((long long)pos * ea.stride[e]) % T);
  float c1, c2;
  if (T > 0) {
    c1 = (float)s / (float)(T - 1);
    c2 = 1.0f - c1;
  } else {
    c1 = 1.0f;
    c2 = 0.0f;
  }
  float qq[kMaxDofDev];
#pragma unroll
  for (int j = 0; j < (NDOF > 0 ? NDOF : kMaxDofDev); ++j)
    if (NDOF > 0 || j < nd) qq[j] = c1 * ea.qs[e + ea.offset * j] + c2 * ea.qe[e + ea.offset * j];
  fk_store<NDOF>(rb, qq, frames, i, fstride);
}

// q of the sample at order position pos of edge e (robotconfigspace.cpp:409-420).
template <int NDOF>
__global__ void __launch_bounds__(kFkBlock)
    fk_edges_kernel(DevRobot rb, EdgeArgs ea, const int* __restrict__ item_edge,
                    const int* __restrict__ item_pos, const int* __restrict__ nitems_dev, int cap,
                    float* __restrict__ frames, long long fstride) {
  int n = *nitems_dev;
  if (n > cap) n = cap;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    fk_edge_item<NDOF>(rb, ea, item_edge, item_pos, i, frames, fstride);
}

__device__ int gcd_i(int a, int b) {
  while (b) {
    int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// testnum (robotconfigspace.cpp:660-673) and the sample order stride: a
// golden-ratio step coprime with testnum, so that positions 0,1,2,... visit
// every sample once, spread along the edge.
template <int BLOCK>
__device__ __forceinline__ void edge_prologue_one(int ndof, float dq, const EdgeArgs& ea, int k,
                                                  int* __restrict__ testnum, int* __restrict__ stride,
                                                  int* __restrict__ max_testnum,
                                                  unsigned long long* __restrict__ total) {
  int T = 0;
  if (k < ea.nedges && ea.skip != nullptr && ea.skip[k]) {
    testnum[k] = 0;
    stride[k] = 1;
    ea.edge_res[k] = kEdgeSkipped;
  } else if (k < ea.nedges) {
    float norm = 0.0f;
    for (int i = 0; i < ndof; ++i) {
      float diff = ea.qe[k + ea.offset * i] - ea.qs[k + ea.offset * i];
      norm += diff * diff;
    }
    norm = __builtin_sqrtf(norm);
    T = (int)(norm / dq) + 1;
    int g = 1;
    if (T > 2) {
      g = (int)((double)T * 0.6180339887498949);
      if (g < 1) g = 1;
      while (gcd_i(g, T) != 1) ++g;
    }
    testnum[k] = T;
    stride[k] = g;
    ea.edge_res[k] = 0;
  }
  // One atomic pair per block: same-address atomics from all CUs serialise
  // at the memory side (about 15 ns each), and one pair per wave took 45 us
  // for 100k edges.  Every thread of the block calls this function (threads
  // past the edges bring T = 0).
  __shared__ int s_max[BLOCK / 64];
  __shared__ unsigned long long s_sum[BLOCK / 64];
  int tmax = T;
  unsigned long long tsum = (unsigned long long)T;
  for (int d = 32; d >= 1; d >>= 1) {
    tmax = max(tmax, __shfl_xor(tmax, d, 64));
    tsum += __shfl_xor(tsum, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s_max[threadIdx.x >> 6] = tmax;
    s_sum[threadIdx.x >> 6] = tsum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; ++w) {
      tmax = max(tmax, s_max[w]);
      tsum += s_sum[w];
    }
    if (tmax > 0) {
      atomicMax(max_testnum, tmax);
      atomicAdd(total, tsum);
    }
  }
}

__global__ void __launch_bounds__(1024)
    edge_prologue_kernel(int ndof, float dq, EdgeArgs ea, int* __restrict__ testnum,
                         int* __restrict__ stride, int* __restrict__ max_testnum,
                         unsigned long long* __restrict__ total) {
  edge_prologue_one<1024>(ndof, dq, ea, blockIdx.x * 1024 + threadIdx.x, testnum, stride, max_testnum, total);
}

// Items for order positions [lo, hi) of every open edge in [e0, e1).
// One atomic per block reserves the block's slots (same-address atomics
// serialise at the memory side; one per wave took 23 us for 100k edges);
// each lane's slots follow by prefix count.  Every thread of the block calls
// it; threads with e >= e1 contribute nothing.
template <int BLOCK>
__device__ __forceinline__ void edge_compact_one(const EdgeArgs& ea, int e, int e1, int lo, int hi,
                                                 int* __restrict__ item_edge, int* __restrict__ item_pos,
                                                 int* __restrict__ nitems) {
  int cnt = 0;
  if (e < e1) {
    int T = ea.testnum[e];
    int open = __hip_atomic_load(&ea.edge_res[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    if (open && T > lo) cnt = (T < hi ? T : hi) - lo;
  }
  // wave-inclusive prefix sum of cnt
  const int lane = threadIdx.x & 63;
  int incl = cnt;
  for (int d = 1; d < 64; d <<= 1) {
    int v = __shfl_up(incl, d, 64);
    if (lane >= d) incl += v;
  }
  __shared__ int s_wave[BLOCK / 64];
  __shared__ int s_base;
  const int w = threadIdx.x >> 6;
  if (lane == 63) s_wave[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int sum = 0;
    for (int i = 0; i < BLOCK / 64; ++i) {
      const int v = s_wave[i];
      s_wave[i] = sum;
      sum += v;
    }
    s_base = sum > 0 ? atomicAdd(nitems, sum) : 0;
  }
  __syncthreads();
  // The wave's runs are consecutive: lane l's items start at the wave's base
  // + its exclusive prefix.  They are written run by run with all 64 lanes,
  // so every store instruction covers consecutive slots.
  const int o = s_base + s_wave[w] + incl - cnt;
  unsigned long long runs = __ballot(cnt > 0);
  while (runs) {
    const int src = __builtin_ctzll(runs);
    runs &= runs - 1;
    const int c = __shfl(cnt, src, 64), os = __shfl(o, src, 64), es = __shfl(e, src, 64);
    for (int t = lane; t < c; t += 64) {
      item_edge[os + t] = es;
      item_pos[os + t] = lo + t;
    }
  }
}

__global__ void __launch_bounds__(1024)
    edge_compact_kernel(EdgeArgs ea, int e0, int e1, int lo, int hi, int* __restrict__ item_edge,
                        int* __restrict__ item_pos, int* __restrict__ nitems) {
  edge_compact_one<1024>(ea, e0 + blockIdx.x * 1024 + threadIdx.x, e1, lo, hi, item_edge, item_pos, nitems);
}

// Multi-scene edge calls: one task per scene (prologue) or per scene part of
// a launch step (compaction, kinematics); block b of a grid belongs to the
// last task whose first block (block0 for the prologue and the compaction,
// fblock0 for the kinematics) is <= b.
template <bool FK>
__device__ __forceinline__ const EdgeTask& edge_task_of(const EdgeTask* __restrict__ tasks, int ntasks) {
  int lo = 0, hi = ntasks - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((FK ? tasks[mid].fblock0 : tasks[mid].block0) <= b) lo = mid;
    else hi = mid - 1;
  }
  return tasks[lo];
}

__global__ void __launch_bounds__(256) edge_prologue_multi_kernel(const EdgeTask* __restrict__ tasks, int ntasks) {
  const EdgeTask& T = edge_task_of<false>(tasks, ntasks);
  edge_prologue_one<256>(T.rb.ndof, T.dq, T.ea, (blockIdx.x - T.block0) * 256 + threadIdx.x, T.testnum, T.stride,
                    T.max_testnum, T.total);
}

__global__ void __launch_bounds__(256) edge_compact_multi_kernel(const EdgeTask* __restrict__ tasks, int ntasks) {
  const EdgeTask& T = edge_task_of<false>(tasks, ntasks);
  edge_compact_one<256>(T.ea, T.e0 + (blockIdx.x - T.block0) * 256 + threadIdx.x, T.e1, T.lo, T.hi, T.item_edge,
                        T.item_pos, T.nitems);
}

template <int NDOF>
__global__ void __launch_bounds__(kFkBlock) fk_edges_multi_kernel(const EdgeTask* __restrict__ tasks, int ntasks) {
  const EdgeTask& T = edge_task_of<true>(tasks, ntasks);
  const int i = (blockIdx.x - T.fblock0) * kFkBlock + threadIdx.x;
  int n = *T.nitems;
  if (n > T.cap) n = T.cap;
  if (i < n) fk_edge_item<NDOF>(T.rb, T.ea, T.item_edge, T.item_pos, i, T.frames, (long long)T.cap);
}

// ---- cell index for the radius search -------------------------------------
// Nodes are binned by their first g.dims coordinates into cells at least D
// wide, so every node within D of a query lies in the 3^dims cells around the
// query's own cell.  The index is a counting sort: per-cell counts, their
// exclusive scan, and a scatter of the coordinates (SoA, stride n) and node
// ids into cell order.


__global__ void grid_count_kernel(CellGrid g, const float* __restrict__ nodes, int n, long long noff,
                                  int* __restrict__ cell_id, int* __restrict__ counts) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  int c = 0;
  for (int d = g.dims - 1; d >= 0; --d) c = c * g.cells[d] + grid_coord(g, d, nodes[l + noff * d]);
  cell_id[l] = c;
  atomicAdd(&counts[c], 1);
}

// Exclusive scan of counts[0, ncells) into starts and cursor (one block of 1024).
__global__ void __launch_bounds__(1024) grid_scan_kernel(const int* __restrict__ counts, int* __restrict__ starts,
                                                         int* __restrict__ cursor, int ncells) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (ncells + 1023) / 1024;
  const int c0 = min(ncells, t * per), c1 = min(ncells, c0 + per);
  int sum = 0;
  for (int c = c0; c < c1; ++c) sum += counts[c];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan of the partial sums
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  for (int c = c0; c < c1; ++c) {
    starts[c] = run;
    cursor[c] = run;
    run += counts[c];
  }
}

__global__ void grid_scatter_kernel(int ndof, const float* __restrict__ nodes, int n, long long noff,
                                    const int* __restrict__ cell_id, int* __restrict__ cursor,
                                    float* __restrict__ sorted, int* __restrict__ sid) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const int pos = atomicAdd(&cursor[cell_id[l]], 1);
  for (int i = 0; i < ndof; ++i) sorted[pos + (size_t)n * i] = nodes[l + noff * i];
  sid[pos] = l;
}

constexpr int pow3(int d) { return d == 0 ? 1 : 3 * pow3(d - 1); }
constexpr int kNeighbourCells = pow3(kGridDims);
static_assert(kNeighbourCells <= 256, "one thread per neighbour cell in a 256-thread block");

// Radius search over the cell index (get_near_vertices,
// prm/src/lib/prm/prmsolver.cpp:2532-2574): one block per query; the nodes
// of the (up to 3^kGridDims) neighbouring cells form one index range that the
// block's threads stride through.  |q_l - query_j|^2 < D^2 in float with the
// reference's summation order; matches (query, node, distance) are appended
// with one atomic per wave.
__global__ void __launch_bounds__(256)
    near_grid_kernel(int ndof, const float* __restrict__ queries, int nq, long long qoff, CellGrid g,
                     const int* __restrict__ starts, const int* __restrict__ counts, const float* __restrict__ sorted,
                     const int* __restrict__ sid, int n, float D2, int* __restrict__ out_q,
                     int* __restrict__ out_node, float* __restrict__ out_d2, int* __restrict__ count, int cap) {
  constexpr int NK = kNeighbourCells;
  __shared__ int s_start[NK], s_pre[NK + 1];
  const int j = blockIdx.x;
  if (j >= nq) return;
  float qr[kMaxDofDev];
  for (int i = 0; i < ndof; ++i) qr[i] = queries[j + qoff * i];
  if (threadIdx.x < NK) {
    // neighbour k: offset digit[d] - 1 in grid dimension d; a dimension
    // without grid (d >= dims) takes only offset 0; a cell whose faces lie
    // D or more from the query is skipped (cell_gap2)
    int digit[kGridDims];
    for (int d = 0, rem = threadIdx.x; d < kGridDims; ++d, rem /= 3) digit[d] = rem % 3;
    int c = 0;
    bool ok = true;
    float lb = 0.0f;
    for (int d = kGridDims - 1; d >= 0; --d) {
      if (d >= g.dims) {
        if (digit[d] != 1) ok = false;
        continue;
      }
      const int x = grid_coord(g, d, qr[d]) + digit[d] - 1;
      if (x < 0 || x >= g.cells[d]) ok = false;
      c = c * g.cells[d] + x;
      lb += cell_gap2(g, d, qr[d], digit[d]);
    }
    if (lb >= D2) ok = false;
    s_start[threadIdx.x] = ok ? starts[c] : 0;
    s_pre[threadIdx.x + 1] = ok ? counts[c] : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_pre[0] = 0;
    for (int k = 0; k < NK; ++k) s_pre[k + 1] += s_pre[k];
  }
  __syncthreads();
  const int total = s_pre[NK];
  const int lane = threadIdx.x & 63;
  for (int base = 0; base < total; base += blockDim.x) {
    const int idx = base + threadIdx.x;
    float normsq = 0.0f;
    bool hit = false;
    int node = 0;
    if (idx < total) {
      // the neighbour cell holding idx: last k with s_pre[k] <= idx
      int k = 0;
      for (int step = 128; step > 0; step >>= 1)
        if (k + step < NK && s_pre[k + step] <= idx) k += step;
      const int pos = s_start[k] + (idx - s_pre[k]);
      node_dist2<1>(sorted, (size_t)n, &pos, ndof, qr, &normsq);
      hit = normsq < D2;
      node = sid[pos];
    }
    unsigned long long m = __ballot(hit);
    if (m == 0ull) continue;
    int b = 0;
    if (lane == 0) b = atomicAdd(count, __popcll(m));
    b = __shfl(b, 0, 64);
    if (hit) {
      int o = b + __popcll(m & ((1ull << lane) - 1ull));
      if (o < cap) {
        out_q[o] = j;
        out_node[o] = node;
        out_d2[o] = normsq;
      }
    }
  }
}

__global__ void fill_i32_kernel(int* p, int v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u32_kernel(unsigned int* p, unsigned int v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ----------------------------------------------------------------------------
// separating-vector kernel
// ----------------------------------------------------------------------------

// Scene image as staged in LDS (layout documented at SceneImg).
struct Lds {
  const float4* verts;          // (x, y, z, info bits): info = row offset << 8 | degree
  const int4* pairs;            // (vbP | vbQ << 16, sysP | sysQ << 8, CRS index, kP | kQ << 16)
  const float4* spheres;        // per polytope, own frame: (centre, sphere radius), (box half extents, |c|_1 + r)
  const unsigned short* adj;    // rows of 4k neighbour ids, padded with the row's own vertex
};

// Per-lane iteration history R[k] (x, y, z) and combs[k] (w: p | q << 16) in
// global memory, entry k of a lane at H[k * hs] (lanes adjacent, so the lanes
// of a wave at the same k touch one contiguous span); only iterations that do
// not separate touch it.

__device__ __forceinline__ int info_of(const float4& r) { return __float_as_int(r.w); }

// support_vertex_searcher::search_support_vertex (collision4.h:66-103):
// first improvement over the stored neighbour order, restart at the new
// vertex.  A vertex's neighbours are read 4 at a time with independent loads;
// the first of them (in stored order) whose value exceeds the current maximum
// is the sequential scan's move, found as the lowest set bit of a compare
// mask.  The new maximum is recomputed from the same operands, so it is the
// same float the scan would keep.
__device__ __forceinline__ int climb(const Lds& L, int p, const V3& dir) {
  float4 rec = L.verts[p];
  float mx = dot3(v3(rec), dir);
  int info = info_of(rec);
  int c = 0;
  for (;;) {
    const int off = info >> 8, cnt = info & 255;
    if (c >= cnt) break;
    const uint2 ids = *reinterpret_cast<const uint2*>(L.adj + off + c);
    const float4 r0 = L.verts[ids.x & 0xffff], r1 = L.verts[ids.x >> 16];
    const float4 r2 = L.verts[ids.y & 0xffff], r3 = L.verts[ids.y >> 16];
    // padding entries are the vertex itself: same operands, same value, never > mx
    unsigned bits = (dot3(v3(r0), dir) > mx) ? 1u : 0u;
    bits |= (dot3(v3(r1), dir) > mx) ? 2u : 0u;
    bits |= (dot3(v3(r2), dir) > mx) ? 4u : 0u;
    bits |= (dot3(v3(r3), dir) > mx) ? 8u : 0u;
    if (bits) {
      const unsigned long long packed = ((unsigned long long)ids.y << 32) | ids.x;
      p = (int)((packed >> (16 * __builtin_ctz(bits))) & 0xffff);
      rec = L.verts[p];
      mx = dot3(v3(rec), dir);
      info = info_of(rec);
      c = 0;
    } else {
      c += 4;
    }
  }
  return p;
}

// orthtrafo23 (collision4.h:111-134)
struct Orth {
  float M0, M1, M2, M3, M4, M5;
};
__device__ __forceinline__ V2 mult23(const Orth& o, const V3& x) {
  return V2{(o.M0 * x.x + o.M1 * x.y) + o.M2 * x.z, (o.M3 * x.x + o.M4 * x.y) + o.M5 * x.z};
}

// Iteration that did not separate (collision4.h:301-323): repetition test,
// w update, half plane (find_half_plane, collision4.h:136-223), reflection.
// Returns a nonzero code when the pair is decided (l+1: no half plane), else
// 0.  Taken by few iterations, but it sets the latency of a full test: the
// history R[0..k-1] lives in memory, so it is read in one pass with
// kHistBatch loads in flight, and the repetition test and the half-plane scan run
// together on each entry.  The scan's outcome counts only when no (p,q)
// repeats, as in the reference, where the repetition test comes first; the
// scan stops changing ra/rb where find_half_plane returns false.
constexpr int kHistBatch = 2;  // history loads in flight (four spill registers at five waves per SIMD)
__device__ __forceinline__ int sva_slow(float4* __restrict__ H, size_t hs, int& k, int l, int pq, const V3& r,
                                        float dp, V3& S, V3& w) {
  H[k * hs] = make_float4(r.x, r.y, r.z, __int_as_float(pq));
  // find_half_plane returns at once (w unchanged) when w.r > 0
  const bool scan = k >= 2 && !(dot3(w, r) > 0);
  Orth o{};
  if (scan) {
    // orthtrafo23 M(rk)
    float x = r.x, y = r.y, z = r.z;
    float d2 = x * x + y * y;
    float d = __builtin_sqrtf(d2);
    o.M0 = x * z / d; o.M1 = y * z / d; o.M2 = -d;
    o.M3 = -y / d;    o.M4 = x / d;     o.M5 = 0.0f;
  }
  bool rep = false, fail = false;
  V3 r0{0.0f, 0.0f, 0.0f};
  V2 ra{0.0f, 0.0f}, rb{0.0f, 0.0f};
  for (int j0 = 0; j0 < k; j0 += kHistBatch) {
    float4 e[kHistBatch];
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u)
      e[u] = j0 + u < k ? H[(j0 + u) * hs] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u) {
      const int j = j0 + u;
      if (j >= k) break;
      rep = rep || __float_as_int(e[u].w) == pq;
      if (j == 0) r0 = v3(e[u]);
      if (!scan || fail) continue;
      const V2 T = mult23(o, v3(e[u]));
      if (j == 0) {
        ra = T;
      } else if (j == 1) {
        rb = T;
        if (cross2(ra, rb) < 0) {
          rb = ra;
          ra = T;
        }
      } else {
        const float rbcross = cross2(rb, T);
        if (cross2(ra, T) > 0) {
          if (rbcross > 0) rb = T;
        } else {
          if (rbcross < 0) ra = T;
          else fail = true;
        }
      }
    }
  }
  if (rep) {
    S = w;  // (p,q) seen before
    return 0;
  }
  if (k == 1) {
    w = V3{r0.x + r.x, r0.y + r.y, r0.z + r.z};
    normalize3(w);
  }
  if (scan) {
    if (fail) return l + 1;
    normalize2(ra);
    normalize2(rb);
    V2 T{ra.x + rb.x, ra.y + rb.y};
    normalize2(T);
    w.x = o.M0 * T.x + o.M3 * T.y;
    w.y = o.M1 * T.x + o.M4 * T.y;
    w.z = o.M2 * T.x + o.M5 * T.y;
  }
  float b = -2.0f * dp;
  S.x = S.x + b * r.x;
  S.y = S.y + b * r.y;
  S.z = S.z + b * r.z;
  ++k;
  return 0;
}

// Work counters of one lane of the separating-vector loop.
struct SvaCounts {
  unsigned sep = 0, col = 0, dropped = 0, units = 0, iter = 0, bound = 0;
  unsigned claimed = 0;  // units this lane claimed (all finished when the loop returns)
};

// Copy a scene image into LDS (every thread of the block takes part; the
// caller synchronises before reading it).
template <int BLOCK>
__device__ __forceinline__ Lds stage_scene(const SceneImg& img, int4* smem) {
  const int n16 = img.bytes >> 4;
  for (int i = threadIdx.x; i < n16; i += BLOCK) smem[i] = img.blob[i];
  const char* base = reinterpret_cast<const char*>(smem);
  Lds L;
  L.verts = reinterpret_cast<const float4*>(base + img.off_verts);
  L.pairs = reinterpret_cast<const int4*>(base + img.off_pairs);
  L.adj = reinterpret_cast<const unsigned short*>(base + img.off_adj);
  L.spheres = reinterpret_cast<const float4*>(base + img.off_spheres);
  return L;
}

// The scene read where it lies in the global image (scenes beyond LDS): the
// same layout, through the caches.
__device__ __forceinline__ Lds scene_global(const SceneImg& img) {
  const char* base = reinterpret_cast<const char*>(img.blob);
  Lds L;
  L.verts = reinterpret_cast<const float4*>(base + img.off_verts);
  L.pairs = reinterpret_cast<const int4*>(base + img.off_pairs);
  L.adj = reinterpret_cast<const unsigned short*>(base + img.off_adj);
  L.spheres = reinterpret_cast<const float4*>(base + img.off_spheres);
  return L;
}

// Deferred pair tests of one wave: a ring of CAP = sva_ring_cap(BLOCK)
// entries in LDS, (item, pair) in ent and the item's edge (edge mode) in
// edge.  Phase B starts as soon as CAP - 64 entries wait, so the ring never
// holds more than CAP - 1.
struct Ring {
  int2* ent;
  int* edge;
};
template <int BLOCK>
__device__ __forceinline__ Ring wave_ring(int4* smem) {
  constexpr int CAP = sva_ring_cap(BLOCK);
  int2* const ent = reinterpret_cast<int2*>(smem);
  int* const edge = reinterpret_cast<int*>(ent + (BLOCK / 64) * CAP);
  const int w = threadIdx.x >> 6;
  return Ring{ent + w * CAP, edge + w * CAP};
}

// Work units of a separating-vector launch: (chunk of pairs_per_unit pairs,
// group of 64 consecutive items), numbered chunk-major.
__device__ __forceinline__ unsigned sva_unit_count(const SvaArgs& A, int nitems) {
  return (unsigned)((nitems + 63) >> 6) * (unsigned)A.units_per_item;
}

// Phase B of sva_units: every lane takes a deferred pair from the wave's
// ring and runs the full separating-vector test (collision4.h:227-332) until
// the ring is empty.  The gate is re-read when a pair is taken.
template <int EDGE, int CAP>
__device__ __forceinline__ void sva_drain(const SvaArgs& A, const Lds& L, const Ring& R,
                                          float4* __restrict__ H, size_t hs, unsigned& d_head, unsigned d_tail,
                                          unsigned& n_sep, unsigned& n_col, unsigned& n_dropped, unsigned& n_iter) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  // the ring's entries were written by other lanes of this wave
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  bool busy = false;
  T12 tp, tq;
#pragma unroll
  for (int i = 0; i < 12; ++i) tp.m[i] = tq.m[i] = 0.0f;
  int bitem = 0, bedge = 0, crs = 0, k = 0, l = 0, p = 0, q = 0, vbP = 0, vbQ = 0;
  bool sysP0 = true, sysQ0 = true;
  V3 S{0, 0, 0}, w{0, 0, 0};
  for (;;) {
    const unsigned long long wm = __ballot(!busy);
    const unsigned avail = d_tail - d_head;
    const unsigned take = (unsigned)__popcll(wm) < avail ? (unsigned)__popcll(wm) : avail;
    if (!busy && (unsigned)__popcll(wm & lt_mask) < take) {
      const unsigned slot = (d_head + (unsigned)__popcll(wm & lt_mask)) & (CAP - 1);
      const int2 e = R.ent[slot];
      bitem = e.x;
      const int4 pd = L.pairs[e.y];
      crs = pd.z;
      // the gate and the frames are loaded together (one memory round
      // trip); the frames of a pair that turns out dropped are not used
      int gv;
      if (EDGE) {
        bedge = R.edge[slot];
        gv = __hip_atomic_load(&A.edge_res[bedge], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        gv = __hip_atomic_load(&A.word[bitem], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      vbP = pd.x & 0xffff;
      vbQ = pd.x >> 16;
      const int sP = pd.y & 255, sQ = pd.y >> 8;
      sysP0 = sP == 0;
      sysQ0 = sQ == 0;
      if (!sysP0) tp = load_frame(A.frames, bitem, A.fstride, sP);
      if (!sysQ0) tq = load_frame(A.frames, bitem, A.fstride, sQ);
      const bool drop = EDGE ? gv != 0 : (gv >> 4) <= crs;
      if (drop) {
        ++n_dropped;
      } else {
        // start of the pair (collision4.h:228-244)
        const V3 vq = v3(L.verts[vbQ]), vp = v3(L.verts[vbP]);
        const V3 Pq = sysQ0 ? vq : t_apply(tq, vq);
        const V3 Pp = sysP0 ? vp : t_apply(tp, vp);
        S = V3{Pq.x - Pp.x, Pq.y - Pp.y, Pq.z - Pp.z};
        const float norm = normalize3(S);
        if (norm <= kEpsBelow) S.x = 1.0f;
        p = vbP;
        q = vbQ;
        k = 0;
        l = 0;
        w = V3{0.0f, 0.0f, 0.0f};
        busy = true;
      }
    }
    d_head += take;
    if (__ballot(busy) == 0ull) {
      if (d_head == d_tail) break;
      continue;
    }
    if (!busy) continue;
    ++n_iter;

    // ---- one iteration (collision4.h:266-300) ------------------------------
    V3 Sp = sysP0 ? S : t_apply_rot_inv(tp, S);
    V3 Sn{-S.x, -S.y, -S.z};
    V3 Sq = sysQ0 ? Sn : t_apply_rot_inv(tq, Sn);
    p = climb(L, p, Sp);
    q = climb(L, q, Sq);
    V3 vp = v3(L.verts[p]), vq = v3(L.verts[q]);
    V3 p0 = sysP0 ? vp : t_apply(tp, vp);
    V3 q0 = sysQ0 ? vq : t_apply(tq, vq);
    V3 r{q0.x - p0.x, q0.y - p0.y, q0.z - p0.z};
    normalize3(r);
    float dp = dot3(S, r);

    bool fin = false;
    int code = 0;
    if (dp >= -kEpsBelow) {
      fin = true;
    } else {
      code = sva_slow(H, hs, k, l, (p - vbP) | ((q - vbQ) << 16), r, dp, S, w);
      if (code != 0) {
        fin = true;
      } else if (++l >= kMaxIter) {
        fin = true;
        code = -1;
      }
    }
    if (fin) {
      if (code != 0) {
        ++n_col;
        if (EDGE) {
          __hip_atomic_store(&A.edge_res[bedge], code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          atomicMin(&A.word[bitem], ((crs + 1) << 4) | (code + 1));
        }
      } else {
        ++n_sep;
      }
      busy = false;
    }
  }
}

// Units of one scene (its image in L) until its queue is exhausted, every
// deferred pair included.  A wave claims a unit (one chunk of pairs, 64
// consecutive items: lane l takes item 64 g + l) and walks the chunk's pairs
// in step, so the pair record, the frame changes and the caches below are
// the same in every lane.  Work alternates between two wave-wide phases:
//   A. for each pair every lane starts the pair on its item (gate, frames,
//      the start direction) and applies the bound test, which decides most
//      pairs; a pair it does not decide is appended to the wave's ring;
//   B. once 64 pairs wait (and at the end), every lane takes deferred pairs
//      and runs the full test until the ring is empty (sva_drain).
// H: this lane's iteration history (stride hs); dq: the wave's ring.  Waves
// run independently: there is no block barrier inside.  cnt.claimed counts
// the wave's units (the same in every lane).
template <int EDGE, int CAP>
__device__ __forceinline__ void sva_units(const SvaArgs& A, const Lds& L, const Ring& R,
                                          float4* __restrict__ H, size_t hs, SvaCounts& cnt) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const int nitems = A.nitems >= 0 ? A.nitems : *A.nitems_dev;
  const int M = A.img.M;
  const int PC = A.pairs_per_unit;
  const unsigned ngroups = (unsigned)((nitems + 63) >> 6);
  const unsigned nunits = sva_unit_count(A, nitems);

  unsigned d_head = 0, d_tail = 0;  // the ring's cursors (the same in every lane)
  unsigned n_sep = 0, n_col = 0, n_dropped = 0, n_units = 0, n_iter = 0, n_bound = 0, n_claimed = 0;

  // Point mode with the codes written here: a unit is complete once the
  // ring holds none of its deferred pairs, which is certain whenever the
  // ring is empty.  Units whose pairs are all walked wait in a per-wave list
  // (the ring's edge array, unused in point mode) until the ring next
  // empties; then each adds one to its group's counter, and the wave that
  // brings a group to units_per_item writes that group's 64 codes.  No wave
  // waits for another.
  const bool fuse = !EDGE && A.group_done != nullptr;
  unsigned n_pend = 0;  // the same in every lane
  auto complete_pending = [&]() {
    for (unsigned k = 0; k < n_pend; ++k) {
      const unsigned cu = (unsigned)R.edge[k];
      const unsigned g = cu % ngroups;
      int old = 0;
      // Words and counters are only touched by agent-scope atomics, which
      // are coherent across the XCDs' L2s; what matters is that this wave's
      // word updates (any lane) have completed before its count, which
      // waiting for all its outstanding memory operations gives.  (An
      // agent-scope release or acquire fence would also write back or
      // invalidate the XCD's whole L2, per wave: it made the launch three
      // times slower.)
      __builtin_amdgcn_s_waitcnt(0);
      if (lane == 0)
        old = __hip_atomic_fetch_add(&A.group_done[g], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      old = __shfl(old, 0, 64);
      if (old == A.units_per_item - 1) {
        // every unit of the group has finished: the words are final (the
        // loads below are issued after the count came back)
        const int it = (int)(g << 6) + lane;
        if (it < nitems) {
          const int w = __hip_atomic_load(&A.word[it], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          A.codes[it] = w == 0x7fffffff ? 0 : (w & 15) - 1;
        }
      }
    }
    n_pend = 0;
    // the list's entries were read by every lane before lane 0 refills it
    __builtin_amdgcn_wave_barrier();
  };

  // Work distribution.  Claims on one shared counter serialise at the
  // memory side, so most units are assigned statically: with static_first,
  // wave w takes units w, w + nwaves, ... for three quarters of the full
  // rounds (at least one round; all units when there are fewer units than
  // waves).  The rest is claimed from
  // the counter in guided chunks of max(1, remaining / (2 nwaves)) units; once
  // fewer units than waves may remain, a wave reads the counter before
  // claiming, so that waves finding the queue drained leave without an
  // atomic.
  const unsigned nwaves = (gridDim.x * blockDim.x) >> 6;
  // wave-uniform values held in scalar registers: the unit index, and from
  // it the pair loop, the pair records and the cache tags below
  const unsigned wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  unsigned nstatic = 0;
  if (A.static_first) {
    const unsigned rounds = nunits / nwaves;
    // fewer units than waves: one each, nothing left to claim
    nstatic = rounds == 0 ? nunits : (rounds * 3 / 4 > 1 ? rounds * 3 / 4 : 1u) * nwaves;
  }
  unsigned su = wave;          // next static unit
  unsigned cur = 0, cend = 0;  // units [cur, cend) claimed dynamically
  unsigned seen = 0;           // the counter as last observed by this wave
  for (;;) {
    unsigned u;
    if (su < nstatic) {
      u = su;
      su += nwaves;
    } else if (cur < cend) {
      u = cur++;
    } else {
      const unsigned left = nunits - nstatic > seen ? nunits - nstatic - seen : 0u;
      if (left == 0) break;
      const unsigned k = left / (2 * nwaves) > 1 ? left / (2 * nwaves) : 1u;
      unsigned b = 0xffffffffu;
      if (lane == 0) {
        bool go = true;
        if (left < nwaves) go = __hip_atomic_load(A.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                                nunits - nstatic;
        if (go) b = atomicAdd(A.counter, k);
      }
      b = __builtin_amdgcn_readfirstlane(b);
      if (b == 0xffffffffu || b >= nunits - nstatic) break;
      seen = b + k;
      cur = nstatic + b;
      cend = cur + k < nunits ? cur + k : nunits;
      u = cur++;
    }
    ++n_claimed;
    const unsigned ch = u / ngroups;
    const int item = (int)((u - ch * ngroups) << 6) + lane;
    const bool active = item < nitems;
    const int p0 = (int)ch * PC, p1 = min(M, p0 + PC);
    // early exit: in edge mode once any sample of the edge is blocked; in
    // point mode for pairs after (in CRS order) a pair known to collide,
    // since the reference returns the first colliding pair's code.  The gate
    // is read lazily so that its memory latency overlaps pair work: the
    // item's word (point mode) or edge index (edge mode) is loaded here, the
    // edge's result two pairs later, and the gate is consulted from the
    // unit's fourth pair on.
    int edge = 0, gate = 0x7fffffff;
    if (active) {
      if (EDGE) edge = A.item_edge[item];
      else gate = __hip_atomic_load(&A.word[item], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 4;
    }
    // The unit's pairs run in segments that end where the ring is drained.
    // The caches are local to a segment, so nothing of them is live across
    // a drain, which has registers of its own.
    int pair = p0;
    while (pair < p1) {
      // wave-uniform caches: the frames of links fsP, fsQ; Q's world bounds for
      // polytope ckQ and its start vertex cvQ
      T12 tp, tq;
#pragma unroll
      for (int i = 0; i < 12; ++i) tp.m[i] = tq.m[i] = 0.0f;
      int fsP = -1, fsQ = -1, ckQ = -1, cvQ = -1;
      V3 cQw{0, 0, 0}, PqW{0, 0, 0};
      // Q's box axes in the world frame, times its half extents
      V3 aQ0{0, 0, 0}, aQ1{0, 0, 0}, aQ2{0, 0, 0};
      float magQ = 0.0f, rQ = 0.0f;
      for (; pair < p1; ++pair) {
        // the record is the same in every lane: keep it in scalar registers
        const int4 pdv = L.pairs[pair];
        const int4 pd = make_int4(__builtin_amdgcn_readfirstlane(pdv.x), __builtin_amdgcn_readfirstlane(pdv.y),
                                  __builtin_amdgcn_readfirstlane(pdv.z), __builtin_amdgcn_readfirstlane(pdv.w));
        const int crs = pd.z;
        const int vbP = pd.x & 0xffff, vbQ = pd.x >> 16;
        const int sP = pd.y & 255, sQ = pd.y >> 8;
        const bool sysP0 = sP == 0, sysQ0 = sQ == 0;
        // the cache tags are wave-uniform: compared as scalars
        // (readfirstlane).  Frames are loaded by every lane: an inactive
        // lane's item is below the plane stride (the item count rounded up
        // to 64), and what it computes from them is masked out below.
        if (!sysP0 && sP != __builtin_amdgcn_readfirstlane(fsP)) {
          tp = load_frame(A.frames, item, A.fstride, sP);
          fsP = sP;
        }
        if (!sysQ0 && sQ != __builtin_amdgcn_readfirstlane(fsQ)) {
          tq = load_frame(A.frames, item, A.fstride, sQ);
          fsQ = sQ;
          ckQ = cvQ = -1;
        }
        const int kP = pd.w & 0xffff, kQ = pd.w >> 16;
        if (vbQ != __builtin_amdgcn_readfirstlane(cvQ)) {
          const V3 vq = v3(L.verts[vbQ]);
          PqW = sysQ0 ? vq : t_apply(tq, vq);
          cvQ = vbQ;
        }
        if (kQ != __builtin_amdgcn_readfirstlane(ckQ)) {
          const float4 b0 = L.spheres[2 * kQ], b1 = L.spheres[2 * kQ + 1];
          cQw = sysQ0 ? v3(b0) : t_apply(tq, v3(b0));
          magQ = fabsf(cQw.x) + fabsf(cQw.y) + fabsf(cQw.z) + b0.w;
          rQ = b0.w;
          if (sysQ0) {
            aQ0 = V3{b1.x, 0.0f, 0.0f};
            aQ1 = V3{0.0f, b1.y, 0.0f};
            aQ2 = V3{0.0f, 0.0f, b1.z};
          } else {
            aQ0 = V3{tq.m[0] * b1.x, tq.m[1] * b1.x, tq.m[2] * b1.x};
            aQ1 = V3{tq.m[3] * b1.y, tq.m[4] * b1.y, tq.m[5] * b1.y};
            aQ2 = V3{tq.m[6] * b1.z, tq.m[7] * b1.z, tq.m[8] * b1.z};
          }
          ckQ = kQ;
        }
        if (EDGE && pair == p0 + 2)  // an inactive lane reads edge 0's result, unused
          gate = __hip_atomic_load(&A.edge_res[edge], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ? 0 : 0x7fffffff;
        // Every lane evaluates the bound test (a SIMD runs the whole wave
        // anyway); the lane's conditions (its item active, its pair not
        // dropped by the gate) are applied to the result, which keeps the
        // exec-mask switching out of the loop.
        // (bitwise & on the conditions: no short-circuit branches)
        const bool drop = active & (pair >= p0 + 4) & (EDGE ? gate == 0 : gate <= crs);
        const bool test = active & !drop;
        // Start of the pair (collision4.h:228-244) and the bound decision
        // of iteration 0: iteration 0 returns 0 whenever
        // S.normalize(q0 - p0) >= -1e-6 for the support vertices p0, q0
        // the climbs find, S = normalize(D) with D the difference of the
        // two start vertices.  If the bounding boxes show S.(q - p) > 0
        // with a wide margin for every vertex p of P and q of Q, that holds
        // whichever vertices the climbs would pick, so the pair test ends
        // here exactly as the reference's does (code 0, one
        // iteration).  The margin, 2^-10 of the magnitudes involved,
        // exceeds the rounding error of every quantity in the test and in
        // the reference's iteration by orders of magnitude.  Both sides of
        // the test scale with |S|, so it runs on D itself: the reference's
        // S differs from D/|D| by a relative 1e-7, far inside the margin.
        // The test also requires |D|_1 > 1e-5, so |D|_2 > 5.7e-6 stays
        // clear of the reference's override of S for |D| <= 1e-6; a NaN D
        // never passes (such pairs are deferred to the exact test).
        const V3 vp = v3(L.verts[vbP]);
        const V3 Pp = sysP0 ? vp : t_apply(tp, vp);
        const V3 S{PqW.x - Pp.x, PqW.y - Pp.y, PqW.z - Pp.z};
        // First the bounding spheres: every vertex lies within r of its
        // polytope's centre c, so S.(q - p) >= S.(cQ - cP) - (rP + rQ)|S|.
        // That costs a few operations once cP is in the world frame and
        // decides most pairs; the boxes' support extents along S (sum_a
        // |S.axis_a| h_a, axes the frame's rotation columns) are tighter
        // and are computed only when some lane's pair the spheres leave
        // open.  The margin is the same in both: 2^-10 of |S|_1 times the
        // magnitudes |c|_1 + r per side (precomputed for world-frame
        // polytopes); the spheres' comparison is made on squares (no square
        // root), both sides non-negative, the right one inflated by 2^-8
        // against their rounding.
        const float4 b0 = L.spheres[2 * kP], b1 = L.spheres[2 * kP + 1];
        const V3 cPw = sysP0 ? v3(b0) : t_apply(tp, v3(b0));
        const float magP = sysP0 ? b1.w : fabsf(cPw.x) + fabsf(cPw.y) + fabsf(cPw.z) + b0.w;
        const float sn1 = fabsf(S.x) + fabsf(S.y) + fabsf(S.z);
        const float dc = dot3(S, V3{cQw.x - cPw.x, cQw.y - cPw.y, cQw.z - cPw.z});
        const float thr = 0x1p-10f * (magQ + magP) * sn1;
        const bool open = (sn1 > 1e-5f) & (dc > thr);
        const float A = dc - thr, rsum = b0.w + rQ;
        bool sep = open & (A * A > (rsum * rsum) * dot3(S, S) * (1.0f + 0x1p-8f));
        const bool need_box = test & open & !sep;
        if (__ballot(need_box) != 0ull) {
          const float boxQ = (fabsf(dot3(S, aQ0)) + fabsf(dot3(S, aQ1))) + fabsf(dot3(S, aQ2));
          float boxP;
          if (sysP0) {
            boxP = (fabsf(S.x) * b1.x + fabsf(S.y) * b1.y) + fabsf(S.z) * b1.z;
          } else {
            const V3 sp = t_apply_rot_inv(tp, S);
            boxP = (fabsf(sp.x) * b1.x + fabsf(sp.y) * b1.y) + fabsf(sp.z) * b1.z;
          }
          sep = sep | (open & (dc - boxQ - boxP > thr));
        }
        sep = sep & test;
        n_dropped += drop ? 1u : 0u;
        n_bound += sep ? 1u : 0u;
        n_sep += sep ? 1u : 0u;
        const bool defer = test & !sep;
        // append the undecided pairs to the ring; drain it once CAP - 64 wait
        const unsigned long long dm = __ballot(defer);
        if (defer) {
          const unsigned slot = (d_tail + (unsigned)__popcll(dm & lt_mask)) & (CAP - 1);
          R.ent[slot] = make_int2(item, pair);
          if (EDGE) R.edge[slot] = edge;
        }
        d_tail += (unsigned)__popcll(dm);
        if (d_tail - d_head >= CAP - 64) {
          ++pair;
          break;
        }
      }
      if (d_tail - d_head >= CAP - 64) {
        sva_drain<EDGE, CAP>(A, L, R, H, hs, d_head, d_tail, n_sep, n_col, n_dropped, n_iter);
        if (fuse) complete_pending();  // the units before this one
      }
    }
    ++n_units;
    if (fuse) {
      if (lane == 0) R.edge[n_pend] = (int)u;
      __builtin_amdgcn_wave_barrier();
      ++n_pend;
      if (d_tail == d_head) {
        complete_pending();
      } else if (n_pend == (unsigned)CAP) {
        sva_drain<EDGE, CAP>(A, L, R, H, hs, d_head, d_tail, n_sep, n_col, n_dropped, n_iter);
        complete_pending();
      }
    }
  }
  if (d_tail != d_head) sva_drain<EDGE, CAP>(A, L, R, H, hs, d_head, d_tail, n_sep, n_col, n_dropped, n_iter);
  if (fuse) complete_pending();

  cnt.sep += n_sep;
  cnt.col += n_col;
  cnt.dropped += n_dropped;
  cnt.units += n_units;
  cnt.iter += n_iter;
  cnt.bound += n_bound;
  cnt.claimed += n_claimed;
}

// Per-block sums of the lane counters through LDS, one atomic per counter.
// Every thread of the block calls it; the LDS scene is not read afterwards.
template <int BLOCK>
__device__ __forceinline__ void add_tally(const SvaCounts& c, int4* smem, unsigned long long* tally) {
  __syncthreads();
  unsigned* t = reinterpret_cast<unsigned*>(smem);
  const int tid = threadIdx.x;
  if (tid < kTallyCount) t[tid] = 0u;
  __syncthreads();
  atomicAdd(&t[0], c.sep);
  atomicAdd(&t[1], c.col);
  atomicAdd(&t[2], c.dropped);
  atomicAdd(&t[3], c.units);
  atomicAdd(&t[4], c.iter);
  atomicAdd(&t[5], c.bound);
  __syncthreads();
  if (tid < kTallyCount)
    atomicAdd(&tally[(blockIdx.x % kTallySlots) * kTallyCount + tid], (unsigned long long)t[tid]);
}

// WAVES: occupancy target in waves per SIMD; the compiler keeps the register
// budget at 512/WAVES VGPRs (5 waves: 96 VGPRs; the edge-mode build spills
// two dwords).
// GS: the scene is read from the global image instead of an LDS copy (for
// scenes whose image does not fit).
template <int EDGE, int BLOCK, int WAVES, int GS>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) sva_kernel(SvaArgs A) {
  extern __shared__ int4 smem[];
  // dynamic LDS: the waves' deferral rings (sva_ring_bytes), then the
  // scene image (unless GS)
  const Lds L = GS ? scene_global(A.img) : stage_scene<BLOCK>(A.img, smem + sva_ring_bytes(BLOCK) / 16);
  const Ring R = wave_ring<BLOCK>(smem);
  __syncthreads();
  float4* __restrict__ H = A.hist + (size_t)blockIdx.x * BLOCK + threadIdx.x;
  const size_t hs = (size_t)gridDim.x * BLOCK;
  SvaCounts c;
  sva_units<EDGE, sva_ring_cap(BLOCK)>(A, L, R, H, hs, c);
  if (A.tally != nullptr) add_tally<BLOCK>(c, smem, A.tally);
}

// Many scenes in one persistent launch.  Blocks start on different scenes
// (block b on scene b mod ntasks) and visit every scene once in rotation:
// a block whose next scene still has unclaimed units stages that scene's
// image into LDS and runs its units; one that finds the queue drained skips
// it.  When a wave leaves a scene, every unit it claimed there is finished;
// the wave whose count completes the scene's total stamps the device clock.
template <int EDGE, int BLOCK, int WAVES>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES, 8)))
    sva_multi_kernel(const SceneTask* __restrict__ tasks, int ntasks, unsigned long long* t_start,
                     unsigned long long* tally) {
  extern __shared__ int4 smem[];
  __shared__ int s_live;
  if (threadIdx.x == 0) atomicMin(t_start, (unsigned long long)wall_clock64());
  float4* __restrict__ H = tasks[0].a.hist + (size_t)blockIdx.x * BLOCK + threadIdx.x;
  const size_t hs = (size_t)gridDim.x * BLOCK;
  // dynamic LDS as in sva_kernel: deferral rings, then the scene image
  const Ring R = wave_ring<BLOCK>(smem);
  SvaCounts c;
  for (int r = 0; r < ntasks; ++r) {
    const int t = (int)((blockIdx.x + (unsigned)r) % (unsigned)ntasks);
    const SvaArgs A = tasks[t].a;
    const int ni = A.nitems >= 0 ? A.nitems : *A.nitems_dev;
    const unsigned nunits = sva_unit_count(A, ni);
    if (threadIdx.x == 0)
      s_live = __hip_atomic_load(A.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nunits ? 1 : 0;
    __syncthreads();
    if (s_live) {
      const Lds L = stage_scene<BLOCK>(A.img, smem + sva_ring_bytes(BLOCK) / 16);
      __syncthreads();
      const unsigned before = c.claimed;
      sva_units<EDGE, sva_ring_cap(BLOCK)>(A, L, R, H, hs, c);
      const unsigned mine = c.claimed - before;  // the wave's units (the same in every lane)
      if ((threadIdx.x & 63) == 0 && mine > 0) {
        const unsigned old = atomicAdd(tasks[t].done, mine);
        if (old + mine == nunits) tasks[t].t_done[0] = (unsigned long long)wall_clock64();
      }
    }
    __syncthreads();  // s_live and the LDS image are reused by the next scene
  }
  if (tally != nullptr) add_tally<BLOCK>(c, smem, tally);
}

// ---- small point-check calls ------------------------------------------------
// One block per configuration of a call of at most kSmallMax: the
// configurations come as kernel arguments and the codes go straight to
// (pinned host) memory, so a call is one launch and one synchronisation.
// Thread 0 computes the frames into LDS; the threads then run the full
// separating-vector test on the pairs t, t + BLOCK, ... (no bound test: it
// only shortcuts pair tests that end separated), skipping pairs later in
// CRS order than a collision already found, and the block keeps the first
// colliding pair's code, as the batched kernels do.

// Frames 1..ndof of one configuration (kinematics.h:27-37), the same
// operations as fk_store.
__device__ void fk_frames(const DevRobot& rb, const float* q, T12* out) {
  T12 T = t_base(rb);
  for (int i = 0; i < rb.ndof; ++i) {
    float qset = rb.q[i], dset = rb.d[i];
    if (rb.types[i] == 0) qset = q[i];
    else dset = q[i];
    const T12 A = t_set_cs(rb.a[i], rb.ca[i], rb.sa[i], qset, dset);
    T = t_lapply(A, T);
    out[i] = T;
  }
}

// The full separating-vector test of one pair (collision4.h:227-332), as
// sva_drain runs it per lane; H: this lane's iteration history, stride hs.
__device__ int sva_pair_full(const Lds& L, const T12& tp, const T12& tq, bool sysP0, bool sysQ0, int vbP, int vbQ,
                             float4* H, size_t hs) {
  const V3 vq0 = v3(L.verts[vbQ]), vp0 = v3(L.verts[vbP]);
  const V3 Pq = sysQ0 ? vq0 : t_apply(tq, vq0);
  const V3 Pp = sysP0 ? vp0 : t_apply(tp, vp0);
  V3 S{Pq.x - Pp.x, Pq.y - Pp.y, Pq.z - Pp.z};
  const float norm = normalize3(S);
  if (norm <= kEpsBelow) S.x = 1.0f;
  int p = vbP, q = vbQ, k = 0;
  V3 w{0.0f, 0.0f, 0.0f};
  for (int l = 0;;) {
    const V3 Sp = sysP0 ? S : t_apply_rot_inv(tp, S);
    const V3 Sn{-S.x, -S.y, -S.z};
    const V3 Sq = sysQ0 ? Sn : t_apply_rot_inv(tq, Sn);
    p = climb(L, p, Sp);
    q = climb(L, q, Sq);
    const V3 vp = v3(L.verts[p]), vq = v3(L.verts[q]);
    const V3 p0 = sysP0 ? vp : t_apply(tp, vp);
    const V3 q0 = sysQ0 ? vq : t_apply(tq, vq);
    V3 r{q0.x - p0.x, q0.y - p0.y, q0.z - p0.z};
    normalize3(r);
    const float dp = dot3(S, r);
    if (dp >= -kEpsBelow) return 0;
    const int code = sva_slow(H, hs, k, l, (p - vbP) | ((q - vbQ) << 16), r, dp, S, w);
    if (code != 0) return code;
    if (++l >= kMaxIter) return -1;
  }
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) point_small_kernel(SceneImg img, DevRobot rb, SmallPoints qs, int* out) {
  extern __shared__ int4 smem[];
  // dynamic LDS: the scene image, the frames, the lanes' iteration history
  const Lds L = stage_scene<BLOCK>(img, smem);
  T12* frames = reinterpret_cast<T12*>(reinterpret_cast<char*>(smem) + img.bytes);
  float4* H = reinterpret_cast<float4*>(reinterpret_cast<char*>(smem) + img.bytes + sizeof(T12) * kMaxDofDev) +
              threadIdx.x;
  __shared__ int s_word;
  const float* q = qs.q[blockIdx.x];
  if (threadIdx.x == 0) {
    // check_boundaries(q) (robotconfigspace.cpp:795-802)
    int ob = 0;
    for (int j = 0; j < rb.ndof; ++j)
      if (q[j] < rb.mins[j] || q[j] > rb.maxs[j]) { ob = 1; break; }
    s_word = ob ? kWordOutOfBounds : kWordFree;
    if (!ob) fk_frames(rb, q, frames);
  }
  __syncthreads();
  if (s_word != kWordOutOfBounds) {
    for (int pair = threadIdx.x; pair < img.M; pair += BLOCK) {
      const int4 pd = L.pairs[pair];
      const int crs = pd.z;
      if ((__hip_atomic_load(&s_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> 4) <= crs) continue;
      const int sP = pd.y & 255, sQ = pd.y >> 8;
      T12 tp, tq;
#pragma unroll
      for (int i = 0; i < 12; ++i) tp.m[i] = tq.m[i] = 0.0f;
      if (sP != 0) tp = frames[sP - 1];
      if (sQ != 0) tq = frames[sQ - 1];
      const int code = sva_pair_full(L, tp, tq, sP == 0, sQ == 0, pd.x & 0xffff, pd.x >> 16, H, BLOCK);
      if (code != 0) atomicMin(&s_word, ((crs + 1) << 4) | (code + 1));
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int w = s_word;
    out[blockIdx.x] = w == kWordFree ? 0 : (w & 15) - 1;
  }
}

// Kinematics of the configurations of many scenes in one launch; block b
// belongs to the task with the largest block0 <= b.
template <int NDOF>
__global__ void __launch_bounds__(kFkBlock) fk_points_multi_kernel(const FkTask* __restrict__ tasks, int ntasks) {
  int lo = 0, hi = ntasks - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const FkTask& T = tasks[lo];
  const int i = (b - T.block0) * kFkBlock + threadIdx.x;
  if (i >= T.n) return;
  const int nd = NDOF > 0 ? NDOF : T.rb.ndof;
  float qq[kMaxDofDev];
#pragma unroll
  for (int j = 0; j < (NDOF > 0 ? NDOF : kMaxDofDev); ++j)
    if (NDOF > 0 || j < nd) qq[j] = T.q[i + T.offset * j];
  int st = 0;
  for (int j = 0; j < nd; ++j)
    if (qq[j] < T.rb.mins[j] || qq[j] > T.rb.maxs[j]) { st = 1; break; }
  T.word[i] = st ? kWordOutOfBounds : kWordFree;
  fk_store<NDOF>(T.rb, qq, T.frames, i, T.fstride);
}

// codes of every task from its result words (same block map).
__global__ void finalize_points_multi_kernel(const FkTask* __restrict__ tasks, int ntasks) {
  int lo = 0, hi = ntasks - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const FkTask& T = tasks[lo];
  const int i = (b - T.block0) * kFkBlock + threadIdx.x;
  if (i >= T.n) return;
  const int w = T.word[i];
  T.codes[i] = w == kWordFree ? 0 : (w & 15) - 1;
}

}  // namespace

// ----------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------

// Dynamic LDS of a separating-vector block: the deferral rings (16 B per
// thread) and the scene image.
size_t sva_lds_bytes(const SceneImg& img, int block, int gscene) {
  return (gscene ? 0 : (size_t)img.bytes) + sva_ring_bytes(block);
}

int sva_max_lds_bytes() { return 160 * 1024; }

#define GPRM_FK_DISPATCH(KERNEL, GRID, BLOCK, ST, ...)                                      \
  switch (rb.ndof) {                                                                        \
    case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;         \
    case 3: hipLaunchKernelGGL(KERNEL<3>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;         \
    case 4: hipLaunchKernelGGL(KERNEL<4>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;         \
    case 5: hipLaunchKernelGGL(KERNEL<5>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;         \
    case 6: hipLaunchKernelGGL(KERNEL<6>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;         \
    case 7: hipLaunchKernelGGL(KERNEL<7>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;         \
    case 8: hipLaunchKernelGGL(KERNEL<8>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;         \
    default: hipLaunchKernelGGL(KERNEL<0>, GRID, BLOCK, 0, ST, __VA_ARGS__); break;        \
  }

// Threads per kinematics block: kFkBlock when that still gives at least two
// blocks per CU, else 256 (65,536 configurations in 1024-thread blocks
// would occupy 64 of the 256 CUs).
static int fk_block(long long n) { return n >= 2LL * 256 * kFkBlock ? kFkBlock : 256; }

hipError_t launch_fk_points(const DevRobot& rb, const float* q, long long offset, int n, float* frames,
                            long long fstride, int* word, int check_bounds, unsigned* zero_counter,
                            hipStream_t st, const unsigned char* skip, int* group_done) {
  if (n <= 0) return hipSuccess;
  const int fb = fk_block(n);
  dim3 grid((n + fb - 1) / fb), block(fb);
  GPRM_FK_DISPATCH(fk_points_kernel, grid, block, st, rb, q, offset, n, frames, fstride, word,
                   check_bounds, zero_counter, skip, group_done);
  return hipGetLastError();
}

hipError_t launch_frames_unpack(const DevRobot& rb, const float* frames, long long fstride, int n, float* out,
                               hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(frames_unpack_kernel, dim3((n + 255) / 256), dim3(256), 0, st, rb, frames, fstride, n, out);
  return hipGetLastError();
}

hipError_t launch_finalize_points(const int* word, int* codes, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(finalize_points_kernel, dim3((n + 255) / 256), dim3(256), 0, st, word, codes, n);
  return hipGetLastError();
}

hipError_t launch_edge_fk(const DevRobot& rb, EdgeArgs ea, const int* item_edge, const int* item_pos,
                          const int* nitems, int cap, float* frames, hipStream_t st) {
  if (cap <= 0) return hipSuccess;
  const int fb = fk_block(cap);
  int blocks = (cap + fb - 1) / fb;
  if (blocks > 4096) blocks = 4096;
  dim3 grid(blocks), block(fb);
  GPRM_FK_DISPATCH(fk_edges_kernel, grid, block, st, rb, ea, item_edge, item_pos, nitems, cap, frames,
                   (long long)cap);
  return hipGetLastError();
}

template <int EDGE, int BLOCK, int WAVES, int GS>
static const void* sva_fn() {
  return reinterpret_cast<const void*>(&sva_kernel<EDGE, BLOCK, WAVES, GS>);
}

template <int GS>
static const void* sva_variant_gs(int edge_mode, int block, int waves) {
  if (block == 256 && waves == 5) return edge_mode ? sva_fn<1, 256, 5, GS>() : sva_fn<0, 256, 5, GS>();
  if (block == 256 && waves == 4) return edge_mode ? sva_fn<1, 256, 4, GS>() : sva_fn<0, 256, 4, GS>();
  if (block == 1024 && waves == 4) return edge_mode ? sva_fn<1, 1024, 4, GS>() : sva_fn<0, 1024, 4, GS>();
  return nullptr;
}

// The kernel variants a shape may name: (block, waves per SIMD), the scene
// in LDS or (gscene) in the global image.
static const void* sva_variant(int edge_mode, int block, int waves, int gscene) {
  return gscene ? sva_variant_gs<1>(edge_mode, block, waves) : sva_variant_gs<0>(edge_mode, block, waves);
}

template <int EDGE, int BLOCK, int WAVES>
static const void* sva_multi_fn() {
  return reinterpret_cast<const void*>(&sva_multi_kernel<EDGE, BLOCK, WAVES>);
}

static const void* sva_multi_variant(int edge_mode, int block, int waves) {
  if (block == 256 && waves == 5) return edge_mode ? sva_multi_fn<1, 256, 5>() : sva_multi_fn<0, 256, 5>();
  if (block == 256 && waves == 4) return edge_mode ? sva_multi_fn<1, 256, 4>() : sva_multi_fn<0, 256, 4>();
  if (block == 1024 && waves == 4) return edge_mode ? sva_multi_fn<1, 1024, 4>() : sva_multi_fn<0, 1024, 4>();
  return nullptr;
}

// Resident blocks over all CUs of a kernel variant with lds dynamic bytes.
static hipError_t resident_grid(const void* const fns[2], int block, size_t lds, int device, int* grid) {
  int cus = 0, blocks = 0;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return e;
  // the occupancy query sees the dynamic LDS only once the attribute allows it
  for (int em = 0; em < 2; ++em) {
    if (!fns[em]) return hipErrorInvalidValue;
    if (lds > 64 * 1024 &&
        (e = hipFuncSetAttribute(fns[em], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess)
      return e;
  }
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fns[1], block, lds);
  if (e != hipSuccess) return e;
  if (blocks < 1) return hipErrorInvalidConfiguration;
  *grid = cus * blocks;
  return hipSuccess;
}

hipError_t sva_multi_shape_finish(size_t img_bytes, SvaShape* sh, int device) {
  const size_t lds = img_bytes + sva_ring_bytes(sh->block);
  if (lds > (size_t)sva_max_lds_bytes() - 16) return hipErrorInvalidValue;
  if (sh->ppu < 1 || sh->ppu > 256) return hipErrorInvalidValue;
  const void* fns[2] = {sva_multi_variant(0, sh->block, sh->waves), sva_multi_variant(1, sh->block, sh->waves)};
  return resident_grid(fns, sh->block, lds, device, &sh->grid);
}

hipError_t sva_shape_finish(const SceneImg& img, SvaShape* sh, int device) {
  sh->gscene = sva_lds_bytes(img, sh->block) > (size_t)sva_max_lds_bytes() ? 1 : 0;
  const size_t lds = sva_lds_bytes(img, sh->block, sh->gscene);
  if (lds > (size_t)sva_max_lds_bytes()) return hipErrorInvalidValue;
  if (sh->ppu < 1 || sh->ppu > 256) return hipErrorInvalidValue;
  const void* fns[2] = {sva_variant(0, sh->block, sh->waves, sh->gscene),
                        sva_variant(1, sh->block, sh->waves, sh->gscene)};
  return resident_grid(fns, sh->block, lds, device, &sh->grid);
}

SvaShape sva_default_shape(const SceneImg& img) {
  SvaShape sh;
  // 256-thread blocks at five waves per SIMD (five blocks per CU) when five
  // scene copies fit in LDS, at four when four fit, else one 1024-thread
  // block sharing a single copy; units of up to 128 pairs (measured fastest
  // on the edge-checking workload; launches shrink them when there are too
  // few items, see unit_pairs in the host runtime)
  if (5 * sva_lds_bytes(img, 256) <= (size_t)sva_max_lds_bytes()) {
    sh.block = 256;
    sh.waves = 5;
  } else if (4 * sva_lds_bytes(img, 256) <= (size_t)sva_max_lds_bytes()) {
    sh.block = 256;
    sh.waves = 4;
  } else if (sva_lds_bytes(img, 1024) <= (size_t)sva_max_lds_bytes()) {
    sh.block = 1024;
    sh.waves = 4;
  } else {
    // beyond LDS: the scene is read from the global image (sva_shape_finish
    // sets gscene), and LDS holds only the rings
    sh.block = 256;
    sh.waves = 5;
  }
  sh.ppu = 128;
  return sh;
}

hipError_t launch_sva(const SvaArgs& a, int edge_mode, const SvaShape& sh, hipStream_t st) {
  const size_t lds = sva_lds_bytes(a.img, sh.block, sh.gscene);
  if (lds > (size_t)sva_max_lds_bytes() || a.hist == nullptr || sh.grid < 1) return hipErrorInvalidValue;
  const void* fn = sva_variant(edge_mode, sh.block, sh.waves, sh.gscene);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<SvaArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(sh.grid), dim3(sh.block), args, lds, st);
}

// LDS of a small point-check block: the image, the frames and BLOCK lanes'
// iteration history.
static size_t point_small_lds(const SceneImg& img, int block) {
  return (size_t)img.bytes + sizeof(T12) * kMaxDofDev + sizeof(float4) * kMaxIter * (size_t)block;
}

hipError_t point_small_prepare(const SceneImg& img) {
  for (int block : {256, 64}) {
    const size_t lds = point_small_lds(img, block);
    if (lds > (size_t)sva_max_lds_bytes()) continue;
    const void* fn = block == 256 ? reinterpret_cast<const void*>(&point_small_kernel<256>)
                                  : reinterpret_cast<const void*>(&point_small_kernel<64>);
    if (lds > 64 * 1024) return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return hipSuccess;
  }
  return hipErrorInvalidValue;
}

hipError_t launch_point_small(const SceneImg& img, const DevRobot& rb, const SmallPoints& qs, int* out,
                              hipStream_t st) {
  if (qs.n <= 0) return hipSuccess;
  if (qs.n > kSmallMax) return hipErrorInvalidValue;
  if (point_small_lds(img, 256) <= (size_t)sva_max_lds_bytes())
    hipLaunchKernelGGL(point_small_kernel<256>, dim3(qs.n), dim3(256), point_small_lds(img, 256), st, img, rb, qs, out);
  else if (point_small_lds(img, 64) <= (size_t)sva_max_lds_bytes())
    hipLaunchKernelGGL(point_small_kernel<64>, dim3(qs.n), dim3(64), point_small_lds(img, 64), st, img, rb, qs, out);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_sva_multi(const SceneTask* d_tasks, int ntasks, int edge_mode, const SvaShape& sh, size_t img_bytes,
                            unsigned long long* t_start, unsigned long long* tally, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  size_t lds = img_bytes + sva_ring_bytes(sh.block);
  if (lds > (size_t)sva_max_lds_bytes() - 16 || sh.grid < 1) return hipErrorInvalidValue;
  const void* fn = sva_multi_variant(edge_mode, sh.block, sh.waves);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<SceneTask**>(&d_tasks), &ntasks, &t_start, &tally};
  return hipLaunchKernel(fn, dim3(sh.grid), dim3(sh.block), args, lds, st);
}

hipError_t launch_fk_points_multi(const FkTask* d_tasks, int ntasks, int nblocks, int ndof, hipStream_t st) {
  if (ntasks <= 0 || nblocks <= 0) return hipSuccess;
  dim3 grid(nblocks), block(kFkBlock);
  switch (ndof) {
    case 2: hipLaunchKernelGGL(fk_points_multi_kernel<2>, grid, block, 0, st, d_tasks, ntasks); break;
    case 3: hipLaunchKernelGGL(fk_points_multi_kernel<3>, grid, block, 0, st, d_tasks, ntasks); break;
    case 4: hipLaunchKernelGGL(fk_points_multi_kernel<4>, grid, block, 0, st, d_tasks, ntasks); break;
    case 5: hipLaunchKernelGGL(fk_points_multi_kernel<5>, grid, block, 0, st, d_tasks, ntasks); break;
    case 6: hipLaunchKernelGGL(fk_points_multi_kernel<6>, grid, block, 0, st, d_tasks, ntasks); break;
    case 7: hipLaunchKernelGGL(fk_points_multi_kernel<7>, grid, block, 0, st, d_tasks, ntasks); break;
    default: hipLaunchKernelGGL(fk_points_multi_kernel<0>, grid, block, 0, st, d_tasks, ntasks); break;
  }
  return hipGetLastError();
}

hipError_t launch_finalize_points_multi(const FkTask* d_tasks, int ntasks, int nblocks, hipStream_t st) {
  if (ntasks <= 0 || nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(finalize_points_multi_kernel, dim3(nblocks), dim3(kFkBlock), 0, st, d_tasks, ntasks);
  return hipGetLastError();
}

hipError_t launch_edge_prologue(const DevRobot& rb, float dq, EdgeArgs ea, int* testnum_out,
                                int* stride_out, int* max_testnum, unsigned long long* total,
                                hipStream_t st) {
  if (ea.nedges <= 0) return hipSuccess;
  dim3 grid((ea.nedges + 1023) / 1024), block(1024);
  hipLaunchKernelGGL(edge_prologue_kernel, grid, block, 0, st, rb.ndof, dq, ea, testnum_out, stride_out,
                     max_testnum, total);
  return hipGetLastError();
}

hipError_t launch_edge_compact(EdgeArgs ea, int e0, int e1, int lo, int hi, int* item_edge,
                               int* item_pos, int* nitems, hipStream_t st) {
  if (e1 <= e0) return hipSuccess;
  dim3 grid((e1 - e0 + 1023) / 1024), block(1024);
  hipLaunchKernelGGL(edge_compact_kernel, grid, block, 0, st, ea, e0, e1, lo, hi, item_edge, item_pos, nitems);
  return hipGetLastError();
}

hipError_t launch_edge_prologue_multi(const EdgeTask* d_tasks, int ntasks, int nblocks, hipStream_t st) {
  if (ntasks <= 0 || nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(edge_prologue_multi_kernel, dim3(nblocks), dim3(256), 0, st, d_tasks, ntasks);
  return hipGetLastError();
}

hipError_t launch_edge_compact_multi(const EdgeTask* d_tasks, int ntasks, int nblocks, hipStream_t st) {
  if (ntasks <= 0 || nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(edge_compact_multi_kernel, dim3(nblocks), dim3(256), 0, st, d_tasks, ntasks);
  return hipGetLastError();
}

hipError_t launch_edge_fk_multi(const EdgeTask* d_tasks, int ntasks, int nblocks, int ndof, hipStream_t st) {
  if (ntasks <= 0 || nblocks <= 0) return hipSuccess;
  dim3 grid(nblocks), block(kFkBlock);
  switch (ndof) {
    case 2: hipLaunchKernelGGL(fk_edges_multi_kernel<2>, grid, block, 0, st, d_tasks, ntasks); break;
    case 3: hipLaunchKernelGGL(fk_edges_multi_kernel<3>, grid, block, 0, st, d_tasks, ntasks); break;
    case 4: hipLaunchKernelGGL(fk_edges_multi_kernel<4>, grid, block, 0, st, d_tasks, ntasks); break;
    case 5: hipLaunchKernelGGL(fk_edges_multi_kernel<5>, grid, block, 0, st, d_tasks, ntasks); break;
    case 6: hipLaunchKernelGGL(fk_edges_multi_kernel<6>, grid, block, 0, st, d_tasks, ntasks); break;
    case 7: hipLaunchKernelGGL(fk_edges_multi_kernel<7>, grid, block, 0, st, d_tasks, ntasks); break;
    default: hipLaunchKernelGGL(fk_edges_multi_kernel<0>, grid, block, 0, st, d_tasks, ntasks); break;
  }
  return hipGetLastError();
}

CellGrid make_cell_grid_dims(int dims, const float* mins, const float* maxs, float D);

CellGrid make_cell_grid(int ndof, const float* mins, const float* maxs, float D) {
  // as many gridded coordinates as kGridMaxCells allows (at most kGridDims)
  for (int dims = ndof < kGridDims ? ndof : kGridDims; dims >= 1; --dims) {
    CellGrid g = make_cell_grid_dims(dims, mins, maxs, D);
    if (g.ncells <= kGridMaxCells || dims == 1) return g;
  }
  return make_cell_grid_dims(1, mins, maxs, D);
}

CellGrid make_cell_grid_dims(int dims, const float* mins, const float* maxs, float D) {
  CellGrid g{};
  g.dims = dims;
  g.ncells = 1;
  for (int d = 0; d < kGridDims; ++d) {
    g.cells[d] = 1;
    g.lo[d] = 0.0f;
    g.inv[d] = 0.0f;
    if (d >= g.dims) continue;
    const float span = maxs[d] - mins[d];
    int c = (span > 0.0f && D > 0.0f) ? (int)std::floor(span / D) : 1;
    c = c < 1 ? 1 : (c > 64 ? 64 : c);
    g.cells[d] = c;
    g.lo[d] = mins[d];
    g.inv[d] = span > 0.0f ? (float)c / span : 0.0f;
    // a cell must stay at least D wide after rounding of inv
    while (g.cells[d] > 1 && (float)(1.0 / g.inv[d]) < D) {
      g.cells[d] -= 1;
      g.inv[d] = (float)g.cells[d] / span;
    }
    g.ncells *= g.cells[d];
  }
  return g;
}

hipError_t launch_grid_build(const CellGrid& g, int ndof, const float* nodes, int n, long long noff, int* cell_id,
                             int* counts, int* starts, int* cursor, float* sorted, int* sid, hipStream_t st) {
  hipError_t e = launch_fill_i32(counts, 0, g.ncells, st);
  if (e != hipSuccess || n <= 0) return e;
  const dim3 grid((n + 255) / 256), block(256);
  hipLaunchKernelGGL(grid_count_kernel, grid, block, 0, st, g, nodes, n, noff, cell_id, counts);
  hipLaunchKernelGGL(grid_scan_kernel, dim3(1), dim3(1024), 0, st, counts, starts, cursor, g.ncells);
  hipLaunchKernelGGL(grid_scatter_kernel, grid, block, 0, st, ndof, nodes, n, noff, cell_id, cursor, sorted, sid);
  return hipGetLastError();
}

hipError_t launch_near_grid(int ndof, const float* queries, int nq, long long qoff, const CellGrid& g,
                            const int* starts, const int* counts, const float* sorted, const int* sid, int n, float D2,
                            int* out_q, int* out_node, float* out_d2, int* count, int cap, hipStream_t st) {
  if (nq <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(near_grid_kernel, dim3(nq), dim3(256), 0, st, ndof, queries, nq, qoff, g, starts, counts, sorted,
                     sid, n, D2, out_q, out_node, out_d2, count, cap);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, int v, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)b), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int* p, unsigned int v, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)b), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

}  // namespace gprm
