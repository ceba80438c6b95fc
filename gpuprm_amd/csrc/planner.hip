// planner.hip -- kernels of the single-process PRM step (prm.cpp).
//
// A step runs on the device from the trees' state to the trees' next state;
// the host only reads back the step's counts (to size the edge checks) and
// the connection flag.  Per step:
//   weights   prefix sums of the base weights 1/(1+surrnum)^3 of both trees
//   sample    candidates around weighted base nodes (get_random_nodes,
//             prmsolver.cpp:2584-2624), then the exact point check
//   select    the first free candidates of each tree, in slot order
//   near      radius search of the new nodes in both trees, grouped by
//             (new node, tree) (find_neighbours / get_near_vertices)
//   edges     [budget ranking], endpoint gather, exact edge checks
//   insert    calc_conn + part3a: nodes with a free edge join the tree with
//             all their free edges; connection detection; cell indices
//
// Integer work (prefix sums, compaction) is written with wave shuffles and
// one LDS array per block; the radius search is one block per (new node,
// tree) group, so every group's edges are contiguous and in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "planner.h"

namespace gprm {
namespace {

constexpr int kScanBlock = 1024;
constexpr int kScanPer = kPlanScanItems / kScanBlock;  // 8 items per thread
constexpr int kMaxWaves = 16;                          // waves of a 1024-thread block

// Counter-based uniform in [0, 1): splitmix64 of key + (ctr + 1) * golden
// ratio, top 53 bits (the same generator as the candidate offsets).
__device__ __forceinline__ double counter_uniform(unsigned long long key, unsigned long long ctr) {
  unsigned long long z = key + (ctr + 1ull) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

// Exclusive scan of v over the block (any multiple of 64 threads up to
// 1024); *total gets the block's sum.  lds holds kMaxWaves + 1 values.
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T* total, T* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  T incl = v;
  for (int d = 1; d < 64; d <<= 1) {
    const T u = __shfl_up(incl, d, 64);
    if (lane >= d) incl += u;
  }
  if (lane == 63) lds[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = 0;
    for (int i = 0; i < nw; ++i) {
      const T x = lds[i];
      lds[i] = s;
      s += x;
    }
    lds[nw] = s;
  }
  __syncthreads();
  const T r = lds[w] + incl - v;
  *total = lds[nw];
  __syncthreads();  // lds is reused by the caller
  return r;
}

__device__ __forceinline__ const PlanTree& tree_of(const PlanTree& t0, const PlanTree& t1, int g) {
  return g ? t1 : t0;
}

// ---- base weights: 3-pass prefix sum over the device node count ----------
__global__ void __launch_bounds__(kScanBlock) plan_weights_partial(PlanTree t0, PlanTree t1) {
  __shared__ double lds[kMaxWaves + 1];
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  const long long n = *T.n;
  const long long base = (long long)blockIdx.x * kPlanScanItems;
  if (base >= n) {
    if (threadIdx.x == 0) T.bsum[blockIdx.x] = 0.0;
    return;
  }
  const long long i0 = base + (long long)threadIdx.x * kScanPer;
  double w[kScanPer];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const long long i = i0 + k;
    double x = 0.0;
    if (i < n) {
      const double a = 1.0 + (double)T.surr[i];
      x = 1.0 / (a * a * a);
    }
    s += x;
    w[k] = s;
  }
  double tot;
  const double ex = block_excl_scan(s, &tot, lds);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (i0 + k < n) T.cum[i0 + k] = ex + w[k];
  if (threadIdx.x == 0) T.bsum[blockIdx.x] = tot;
}

// Exclusive prefix of the block sums in place, total at [nb] (one block per tree).
template <class V>
__device__ __forceinline__ void scan_block_sums(V* bsum, int nb) {
  __shared__ V lds[kMaxWaves + 1];
  const int per = (nb + kScanBlock - 1) / kScanBlock;
  const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  V s = 0;
  for (int b = b0; b < b1; ++b) s += bsum[b];
  V tot;
  V run = block_excl_scan(s, &tot, lds);
  for (int b = b0; b < b1; ++b) {
    const V x = bsum[b];
    bsum[b] = run;
    run += x;
  }
  if (threadIdx.x == 0) bsum[nb] = tot;
}

__global__ void __launch_bounds__(kScanBlock) plan_weights_top(PlanTree t0, PlanTree t1, int nb) {
  scan_block_sums<double>(tree_of(t0, t1, blockIdx.x).bsum, nb);
}

__global__ void __launch_bounds__(kScanBlock) plan_weights_add(PlanTree t0, PlanTree t1) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  if (blockIdx.x == 0) return;
  const long long n = *T.n;
  const long long base = (long long)blockIdx.x * kPlanScanItems;
  if (base >= n) return;
  const double add = T.bsum[blockIdx.x];
  for (long long i = base + threadIdx.x; i < base + kPlanScanItems && i < n; i += kScanBlock) T.cum[i] += add;
}

// ---- candidates -------------------------------------------------------------
__global__ void __launch_bounds__(256)
    plan_sample_kernel(int ndof, int batch, PlanTree t0, PlanTree t1, int cap, int nb, float D,
                       unsigned long long key, float* __restrict__ cand) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= batch) return;
  const PlanTree& T = tree_of(t0, t1, k & 1);
  const int n = *T.n;
  // base node: the first m with cum[m] > u * total
  const double x = counter_uniform(key ^ 0xd1b54a32d192ed03ull, (unsigned long long)k) * T.bsum[nb];
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (T.cum[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  const int m = lo;
  for (int i = 0; i < ndof; ++i) {
    const float u = (float)counter_uniform(key, (unsigned long long)k * ndof + i);
    cand[k + (size_t)batch * i] = T.q[m + (size_t)cap * i] - D + 2 * D * u;
  }
}

__global__ void __launch_bounds__(256) plan_bias_mask_kernel(const float* __restrict__ prob, int batch,
                                                             unsigned char* __restrict__ skip) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < batch) skip[k] = ((k & 2) && prob[k] > 0.5f) ? 1 : 0;
}

// New nodes in slot order: at most want[g] free, unskipped candidates of
// tree g (slot k is of tree k & 1).  Tiles of 1024 slots, three passes:
// per-tile counts, their prefix (one block), then every taken candidate
// writes itself at j = min(r0, want0) + min(r1, want1), r_g the accepted
// candidates of tree g before it.
__device__ __forceinline__ bool accepted_at(const int* codes, const unsigned char* skip, int k) {
  return !(skip && skip[k]) && codes[k] == 0;
}

__global__ void __launch_bounds__(kScanBlock)
    plan_select_count(int batch, const int* __restrict__ codes, const unsigned char* __restrict__ skip,
                      int* __restrict__ tile) {
  __shared__ int lds[kMaxWaves + 1];
  const int k = blockIdx.x * kScanBlock + threadIdx.x;
  const bool in = k < batch;
  const bool acc = in && accepted_at(codes, skip, k);
  int A0, A1, DR;
  (void)block_excl_scan((int)(acc && !(k & 1)), &A0, lds);
  (void)block_excl_scan((int)(acc && (k & 1)), &A1, lds);
  (void)block_excl_scan((int)(in && skip && skip[k]), &DR, lds);
  if (threadIdx.x == 0) {
    tile[3 * blockIdx.x] = A0;
    tile[3 * blockIdx.x + 1] = A1;
    tile[3 * blockIdx.x + 2] = DR;
  }
}

// Exclusive prefix of the tiles' counts in place; totals into info/state.
__global__ void __launch_bounds__(kScanBlock)
    plan_select_prefix(int batch, int ntiles, int want0, int want1, int* __restrict__ tile, StepInfo* info,
                       PlanState* state) {
  __shared__ int lds[kMaxWaves + 1];
  const int per = (ntiles + kScanBlock - 1) / kScanBlock;
  const int t0 = threadIdx.x * per, t1 = min(ntiles, t0 + per);
  int s[3] = {0, 0, 0}, tot[3];
  for (int t = t0; t < t1; ++t)
    for (int c = 0; c < 3; ++c) s[c] += tile[3 * t + c];
  int run[3];
  for (int c = 0; c < 3; ++c) run[c] = block_excl_scan(s[c], &tot[c], lds);
  for (int t = t0; t < t1; ++t)
    for (int c = 0; c < 3; ++c) {
      const int x = tile[3 * t + c];
      tile[3 * t + c] = run[c];
      run[c] += x;
    }
  if (threadIdx.x == 0) {
    info->nnew = min(tot[0], want0) + min(tot[1], want1);
    info->accepted = tot[0] + tot[1];
    info->E_all = 0;  // the radius search's edge counter (its appending pass)
    atomicAdd(&state->point_checks, (unsigned long long)(batch - tot[2]));
    atomicAdd(&state->dropped, (unsigned long long)tot[2]);
  }
}

__global__ void __launch_bounds__(kScanBlock)
    plan_select_write(int ndof, int batch, const int* __restrict__ codes, const unsigned char* __restrict__ skip,
                      const float* __restrict__ cand, int want0, int want1, int num, const int* __restrict__ tile,
                      float* __restrict__ qnew) {
  __shared__ int lds[kMaxWaves + 1];
  const int k = blockIdx.x * kScanBlock + threadIdx.x;
  const bool acc = k < batch && accepted_at(codes, skip, k);
  int T0, T1;
  const int r0 = tile[3 * blockIdx.x] + block_excl_scan((int)(acc && !(k & 1)), &T0, lds);
  const int r1 = tile[3 * blockIdx.x + 1] + block_excl_scan((int)(acc && (k & 1)), &T1, lds);
  if (!acc || ((k & 1) ? r1 >= want1 : r0 >= want0)) return;
  const int j = min(r0, want0) + min(r1, want1);
  for (int i = 0; i < ndof; ++i) qnew[j + (size_t)num * i] = cand[k + (size_t)batch * i];
}

// ---- radius search -----------------------------------------------------------

constexpr int pow3(int d) { return d == 0 ? 1 : 3 * pow3(d - 1); }
constexpr int kNeighbourCells = pow3(kGridDims);
static_assert(kNeighbourCells <= 256, "one thread per neighbour cell in a 256-thread block");

// One block per group b = 2 j + g: the nonempty cells among the 3^dims
// around new node j are compacted (block scans) and the block strides
// through their nodes in cell order, 512 candidates a chunk (two a thread,
// their loads in flight together).  In either mode the matches are staged
// in LDS (per-wave ballots, one barrier a chunk), so no chunk waits on a
// global atomic.
//
// APPEND (no edge budget, so no group order needed): the group takes one
// range of info->E_all for its staged matches at the end, and one more each
// time the stage fills before that; the first nbuf slots are stored, cnt[b]
// gets the group's stored edges.
//
// GROUP (edge budget: the ranking needs each group's edges contiguous):
// the matches are staged with their traversal index; a group keeps at most
// `keep` of them, its shortest (ties by traversal index),
// keep = min(kGroupKeep, nbuf / groups),
// so a step never has more than nbuf and the kept edges do not depend on
// scheduling.  When the stage would overflow, and at the end when more than
// keep are staged, a bitonic sort in LDS orders the stage and cuts it to
// keep.  At the end the block takes one range of info->E_all and copies its
// edges out: group b's are contiguous at off[b] (in traversal order, or
// sorted when the group was cut).
enum NearMode { kNearAppend = 0, kNearGroup = 1 };
// 12 KB of stage: eight blocks of four waves still fit a CU's LDS (the
// search is latency-bound, so occupancy counts)
constexpr int kGroupStage = 1024;
constexpr int kGroupKeep = kGroupStage / 2;
// Two candidates per thread and chunk (twice the loads in flight, half the
// barriers).  The stage takes a cut group and one chunk, or two chunks.
constexpr int kNearPer = 2;
static_assert(kGroupStage >= kGroupKeep + kNearPer * 256 && kGroupStage >= 2 * kNearPer * 256, "stage");

// Ascending bitonic sort of the first n stage entries by (squared distance,
// traversal index), padded to a power of two with +inf; n <= kGroupStage.
__device__ void stage_sort(float* d2, int* idx, int* node, int n) {
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = n + (int)threadIdx.x; i < P; i += blockDim.x) {
    d2[i] = __builtin_inff();
    idx[i] = 0x7fffffff;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < P / 2; t += blockDim.x) {
        const int a = ((t & ~(j - 1)) << 1) | (t & (j - 1)), b = a | j;  // a has bit j clear
        const float da = d2[a], db = d2[b];
        const int ia = idx[a], ib = idx[b];
        const bool gt = da > db || (da == db && ia > ib);
        if (gt == ((a & k) == 0)) {
          d2[a] = db;
          d2[b] = da;
          idx[a] = ib;
          idx[b] = ia;
          const int na = node[a];
          node[a] = node[b];
          node[b] = na;
        }
      }
      __syncthreads();
    }
}

template <int MODE>
__global__ void __launch_bounds__(256)
    plan_near_kernel(int ndof, int num, const float* __restrict__ qnew, StepInfo* __restrict__ info,
                     PlanTree t0, PlanTree t1, int cap, CellGrid g, float D2, int* __restrict__ cnt,
                     int* __restrict__ off, int nbuf, int keep, int* __restrict__ enode,
                     float* __restrict__ ed2, int* __restrict__ egroup) {
  constexpr bool GROUP = MODE == kNearGroup;
  constexpr int PER = kNearPer;
  constexpr int NK = kNeighbourCells;
  constexpr int kWaves = 256 / 64;
  __shared__ int s_start[NK], s_pre[NK + 1];
  __shared__ int lds[kMaxWaves + 1];
  __shared__ int s_base;
  __shared__ int s_wcnt[2][PER][kWaves];
  __shared__ int s_node[kGroupStage];
  __shared__ float s_d2[kGroupStage];
  __shared__ int s_idx[GROUP ? kGroupStage : 1];
  const int b = blockIdx.x, j = b >> 1;
  if (j >= info->nnew) {
    if (threadIdx.x == 0) cnt[b] = 0;
    return;
  }
  int chunk_no = 0;  // chunks done (the same in every thread)
  const PlanTree& T = tree_of(t0, t1, b & 1);
  float qr[kMaxDofDev];
  for (int i = 0; i < ndof; ++i) qr[i] = qnew[j + (size_t)num * i];
  int cstart = 0, ccount = 0;
  if (threadIdx.x < NK) {
    // neighbour k: offset digit(d) - 1 in grid dimension d; a dimension
    // without grid (d >= dims) takes only offset 0.  A cell whose faces lie
    // D or more from the query is skipped (on a grid of cells D wide about
    // half of the 3^5).
    int rem = threadIdx.x, c = 0;
    bool ok = true;
    int digit[kGridDims];
    for (int d = 0; d < kGridDims; ++d) {
      digit[d] = rem % 3;
      rem /= 3;
    }
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
    if (ok) {
      cstart = T.starts[c];
      ccount = T.counts[c];
    }
  }
  int K, total;
  const int slot = block_excl_scan((int)(ccount > 0), &K, lds);
  const int pre = block_excl_scan(ccount, &total, lds);
  if (ccount > 0) {
    s_start[slot] = cstart;
    s_pre[slot] = pre;
  }
  if (threadIdx.x == 0) s_pre[K] = total;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  // position in the tree's cell order of candidate idx < total of the
  // group's traversal
  auto cell_pos = [&](int idx) -> int {
    int k = 0;  // the nonempty cell holding idx: last k with s_pre[k] <= idx
    for (int step = 128; step > 0; step >>= 1)
      if (k + step < K && s_pre[k + step] <= idx) k += step;
    return s_start[k] + (idx - s_pre[k]);
  };
  // matches of the chunk's u-th candidates in the waves before this one
  // (tot: in the block), from each wave's counts cw: the counts go to one of
  // two LDS rows (alternating, so a row is rewritten only after every thread
  // has passed the next chunk's barrier)
  auto waves_before = [&](const int* cw, int* before, int* tot) {
    int(*const row)[kWaves] = s_wcnt[chunk_no & 1];
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < PER; ++u) row[u][wv] = cw[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      before[u] = 0;
      tot[u] = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const int c = row[u][w];
        before[u] += w < wv ? c : 0;
        tot[u] += c;
      }
    }
    ++chunk_no;
  };
  // matches so far, staged entries, entries stored by earlier flushes
  // (APPEND); the same in every thread
  int run = 0, n = 0, stored = 0;
  for (int base = 0; base < total; base += PER * blockDim.x) {
    if (n + PER * (int)blockDim.x > kGroupStage) {
      if (GROUP) {
        stage_sort(s_d2, s_idx, s_node, n);
        n = min(n, keep);
      } else {
        __syncthreads();  // the last chunk's entries
        if (threadIdx.x == 0) s_base = atomicAdd(&info->E_all, n);
        __syncthreads();
        const int o = s_base;
        const int kept = max(0, min(n, nbuf - o));
        for (int r = threadIdx.x; r < kept; r += blockDim.x) {
          enode[o + r] = s_node[r];
          ed2[o + r] = s_d2[r];
          egroup[o + r] = b;
        }
        stored += kept;
        n = 0;
        __syncthreads();  // the stage is read before it is refilled
      }
    }
    // candidates base + u 256 + t for u < PER: their loads are in flight
    // together (a lane past the end reads the last candidate and matches
    // nothing); the chunk's matches are staged in traversal order
    int pos[PER], node[PER];
    float normsq[PER];
    bool hit[PER];
    unsigned long long m[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) pos[u] = cell_pos(min(base + u * (int)blockDim.x + (int)threadIdx.x, total - 1));
#pragma unroll
    for (int u = 0; u < PER; ++u) node[u] = T.sid[pos[u]];
    node_dist2<PER>(T.sorted, (size_t)cap, pos, ndof, qr, normsq);
    int cw[PER], before[PER], tot[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      hit[u] = base + u * (int)blockDim.x + (int)threadIdx.x < total && normsq[u] < D2;
      m[u] = __ballot(hit[u]);
      cw[u] = __popcll(m[u]);
    }
    waves_before(cw, before, tot);
    int r = 0;  // the chunk's matches among earlier candidates u
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      if (hit[u]) {
        const int rank = r + before[u] + __popcll(m[u] & below);
        s_node[n + rank] = node[u];
        s_d2[n + rank] = normsq[u];
        if (GROUP) s_idx[n + rank] = run + rank;
      }
      r += tot[u];
    }
    run += r;
    n += r;
  }
  if (GROUP && n > keep) {
    __syncthreads();
    stage_sort(s_d2, s_idx, s_node, n);
    n = keep;
  }
  if (threadIdx.x == 0) s_base = n > 0 ? atomicAdd(&info->E_all, n) : 0;
  __syncthreads();
  const int o = s_base;
  // GROUP: n <= keep <= nbuf / groups, never cut
  const int kept = max(0, min(n, nbuf - o));
  if (threadIdx.x == 0) {
    cnt[b] = stored + kept;
    if (GROUP) off[b] = min(o, nbuf);
  }
  for (int r = threadIdx.x; r < kept; r += blockDim.x) {
    enode[o + r] = s_node[r];
    ed2[o + r] = s_d2[r];
    egroup[o + r] = b;
  }
}

// After the radius search: info->E = min(E_all, nbuf); with budget B > 0
// sc_off = exclusive scan over the groups of their stored edges where they
// are more than B (the groups the model scores), info->S its total.
__global__ void __launch_bounds__(kScanBlock)
    plan_near_done_kernel(int groups, const int* __restrict__ cnt, int nbuf, int budget, int* __restrict__ sc_off,
                          StepInfo* info) {
  __shared__ int lds[kMaxWaves + 1];
  int S = 0;
  if (budget > 0) {
    const int per = (groups + kScanBlock - 1) / kScanBlock;
    const int b0 = threadIdx.x * per, b1 = min(groups, b0 + per);
    int sc = 0;
    for (int b = b0; b < b1; ++b) sc += cnt[b] > budget ? cnt[b] : 0;
    int srun = block_excl_scan(sc, &S, lds);
    for (int b = b0; b < b1; ++b) {
      sc_off[b] = srun;
      srun += cnt[b] > budget ? cnt[b] : 0;
    }
  }
  if (threadIdx.x == 0) {
    const int all = info->E_all;
    info->E = all < nbuf ? all : nbuf;
    info->S = S;
    if (budget > 0) sc_off[groups] = S;
  }
}

// Stored edges of group b.
__device__ __forceinline__ int stored_edges(const int* off, const int* cnt, int b, int nbuf) {
  (void)off;
  (void)nbuf;
  return cnt[b];
}

// ---- edge budget ----------------------------------------------------------------
__global__ void __launch_bounds__(256)
    plan_score_points_kernel(int ndof, int E, int nbuf, int budget, int S8, const int* __restrict__ cnt,
                             const int* __restrict__ off, const int* __restrict__ sc_off,
                             const int* __restrict__ egroup, const int* __restrict__ enode,
                             const float* __restrict__ qnew, int num, PlanTree t0, PlanTree t1, int cap,
                             float* __restrict__ out, long long ns) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int b = egroup[e];
  if (stored_edges(off, cnt, b, nbuf) <= budget) return;
  const long long p = sc_off[b] + (e - off[b]);
  const int j = b >> 1, node = enode[e];
  const PlanTree& T = tree_of(t0, t1, b & 1);
  for (int s = 0; s < S8; ++s) {
    const float t = (float)(s + 1) / (float)(S8 + 1);
    for (int i = 0; i < ndof; ++i) {
      const float qa = qnew[j + (size_t)num * i], qb = T.q[node + (size_t)cap * i];
      out[p * S8 + s + ns * i] = qa + t * (qb - qa);
    }
  }
}

__global__ void __launch_bounds__(256)
    plan_edge_score_kernel(int E, int nbuf, int budget, int S8, const int* __restrict__ cnt,
                           const int* __restrict__ off, const int* __restrict__ sc_off,
                           const int* __restrict__ egroup, const float* __restrict__ prob, float* __restrict__ score) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int b = egroup[e];
  float m = 0.0f;
  if (prob && stored_edges(off, cnt, b, nbuf) > budget) {
    const long long p = sc_off[b] + (e - off[b]);
    for (int s = 0; s < S8; ++s) m = fmaxf(m, prob[p * S8 + s]);
  }
  score[e] = m;
}

__global__ void __launch_bounds__(256)
    plan_rank_kernel(int E, int nbuf, int budget, const int* __restrict__ cnt, const int* __restrict__ off,
                     const int* __restrict__ egroup, const float* __restrict__ ed2, const float* __restrict__ score,
                     unsigned char* __restrict__ skip, PlanState* state) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  bool sk = false;
  if (e < E) {
    const int b = egroup[e];
    const int w = stored_edges(off, cnt, b, nbuf);
    if (w > budget) {
      // key: (predicted blocked, squared length, index); predicted blocked
      // means a score above one half
      const int o = off[b];
      const bool be = score[e] > 0.5f;
      const float de = ed2[e];
      // skipped once budget edges rank ahead: only the kept edges scan
      // the whole group
      int rank = 0;
      for (int x = o; x < o + w && rank < budget; ++x) {
        const bool bx = score[x] > 0.5f;
        const float dx = ed2[x];
        rank += (bx < be) || (bx == be && (dx < de || (dx == de && x < e)));
      }
      sk = rank >= budget;
    }
    skip[e] = sk ? 1 : 0;
  }
  const unsigned long long m = __ballot(sk);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&state->edges_skipped, (unsigned long long)__popcll(m));
}

__global__ void __launch_bounds__(256)
    plan_gather_kernel(int ndof, int E, const int* __restrict__ egroup, const int* __restrict__ enode,
                       const float* __restrict__ qnew, int num, PlanTree t0, PlanTree t1, int cap,
                       float* __restrict__ qs, float* __restrict__ qe) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int b = egroup[e], j = b >> 1, node = enode[e];
  const PlanTree& T = tree_of(t0, t1, b & 1);
  // Canonical orientation: the lexicographically smaller endpoint first.
  // The samples of an edge depend on its direction at the last bit
  // (q = c1 qs + c2 qe, robotconfigspace.cpp:409-420), so a fixed orientation
  // lets anyone re-check a tree edge or a path edge exactly as it was
  // decided, from its two configurations alone.
  int order = 0;  // -1: new node first, 1: tree node first
  for (int i = 0; i < ndof && order == 0; ++i) {
    const float a = qnew[j + (size_t)num * i], c = T.q[node + (size_t)cap * i];
    order = a < c ? -1 : (c < a ? 1 : 0);
  }
  for (int i = 0; i < ndof; ++i) {
    const float a = qnew[j + (size_t)num * i], c = T.q[node + (size_t)cap * i];
    qs[e + (size_t)E * i] = order > 0 ? c : a;
    qe[e + (size_t)E * i] = order > 0 ? a : c;
  }
}

__global__ void __launch_bounds__(256)
    plan_edges_done_kernel(int E, const int* __restrict__ egroup, const int* __restrict__ res,
                           const unsigned char* __restrict__ skip, int* __restrict__ free_cnt,
                           const unsigned long long* __restrict__ total, const int* __restrict__ max_t,
                           int maxT_bound, PlanState* state) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  bool checked = false, fr = false;
  if (e < E) {
    checked = !(skip && skip[e]);
    fr = res[e] == 0;  // skipped edges carry a nonzero result
    if (fr) atomicAdd(&free_cnt[egroup[e]], 1);
  }
  const unsigned long long mc = __ballot(checked), mf = __ballot(fr);
  if ((threadIdx.x & 63) == 0) {
    if (mc) atomicAdd(&state->edges_checked, (unsigned long long)__popcll(mc));
    if (mf) atomicAdd(&state->edges_free, (unsigned long long)__popcll(mf));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(&state->edge_samples, *total);
    if (*max_t > maxT_bound) state->over_bound = 1;
  }
}

// ---- insertion ---------------------------------------------------------------------
__global__ void __launch_bounds__(kScanBlock)
    plan_insert_plan_kernel(int num, const StepInfo* __restrict__ info, const int* __restrict__ free_cnt,
                            int* __restrict__ pos, int cap, PlanState* state) {
  __shared__ int lds[kMaxWaves + 1];
  __shared__ int s_conn;
  const int nnew = info->nnew;
  const int n0 = state->n[0], n1 = state->n[1];
  const bool searching = state->conn == 0;
  if (threadIdx.x == 0) s_conn = 0x7fffffff;
  __syncthreads();
  const int per = (num + kScanBlock - 1) / kScanBlock;
  const int j0 = threadIdx.x * per, j1 = min(num, j0 + per);
  // the first new node with free edges to both trees
  if (searching) {
    for (int j = j0; j < j1 && j < nnew; ++j)
      if (free_cnt[2 * j] > 0 && free_cnt[2 * j + 1] > 0) {
        atomicMin(&s_conn, j);
        break;
      }
  }
  __syncthreads();
  const int cj = s_conn;
  const int jmax = cj != 0x7fffffff ? cj : nnew - 1;
  int a0 = 0, a1 = 0;
  for (int j = j0; j < j1 && j <= jmax; ++j) {
    a0 += free_cnt[2 * j] > 0;
    a1 += free_cnt[2 * j + 1] > 0;
  }
  int A0, A1;
  int r0 = block_excl_scan(a0, &A0, lds);
  int r1 = block_excl_scan(a1, &A1, lds);
  for (int j = j0; j < j1; ++j) {
    int p0 = -1, p1 = -1;
    if (j <= jmax) {
      if (free_cnt[2 * j] > 0) {
        if (n0 + r0 < cap) p0 = n0 + r0;
        ++r0;
      }
      if (free_cnt[2 * j + 1] > 0) {
        if (n1 + r1 < cap) p1 = n1 + r1;
        ++r1;
      }
    }
    pos[2 * j] = p0;
    pos[2 * j + 1] = p1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const bool full = n0 + A0 > cap || n1 + A1 > cap;
    if (full) state->full = 1;
    state->n[0] = min(cap, n0 + A0);
    state->n[1] = min(cap, n1 + A1);
    if (cj != 0x7fffffff && pos[2 * cj] >= 0 && pos[2 * cj + 1] >= 0) {
      state->conn = 1;
      state->conn_left = pos[2 * cj];
      state->conn_right = pos[2 * cj + 1];
    }
  }
}

__global__ void __launch_bounds__(256)
    plan_insert_nodes_kernel(int ndof, int num, const int* __restrict__ pos, const int* __restrict__ free_cnt,
                             const float* __restrict__ qnew, PlanTree t0, PlanTree t1, int cap) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= 2 * num) return;
  const int p = pos[b];
  if (p < 0) return;
  const PlanTree& T = tree_of(t0, t1, b & 1);
  const int j = b >> 1;
  for (int i = 0; i < ndof; ++i) T.q[p + (size_t)cap * i] = qnew[j + (size_t)num * i];
  T.surr[p] = free_cnt[b];
}

__global__ void __launch_bounds__(256)
    plan_insert_edges_kernel(int E, const int* __restrict__ egroup, const int* __restrict__ enode,
                             const float* __restrict__ ed2, const int* __restrict__ res, const int* __restrict__ pos,
                             PlanTree t0, PlanTree t1, PlanState* state) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  int b = 0, p = -1;
  if (e < E && res[e] == 0) {
    b = egroup[e];
    p = pos[b];
  }
  const int g = b & 1;
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1ull;
  // one edge-counter atomic per wave and tree
  unsigned x = 0;
  for (int t = 0; t < 2; ++t) {
    const unsigned long long m = __ballot(p >= 0 && g == t);
    if (m == 0ull) continue;
    const int leader = __builtin_ctzll(m);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(tree_of(t0, t1, t).ecount, (unsigned)__popcll(m));
    base = __shfl(base, leader, 64);
    if (p >= 0 && g == t) x = base + (unsigned)__popcll(m & lt);
  }
  if (p < 0) return;
  const PlanTree& T = tree_of(t0, t1, g);
  const int node = enode[e];
  atomicAdd(&T.surr[node], 1);
  if (x < (unsigned)T.ecap) {
    T.efrom[x] = p;
    T.eto[x] = node;
    T.ew[x] = sqrtf(ed2[e]);
  } else {
    state->full = 1;
  }
}

// ---- cell index ----------------------------------------------------------------------
__global__ void __launch_bounds__(256) plan_fill_counts(PlanTree t0, PlanTree t1, int ncells) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) T.counts[c] = 0;
}

__global__ void __launch_bounds__(256) plan_grid_count(PlanTree t0, PlanTree t1, int cap, CellGrid g) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= *T.n) return;
  int c = 0;
  for (int d = g.dims - 1; d >= 0; --d) c = c * g.cells[d] + grid_coord(g, d, T.q[l + (size_t)cap * d]);
  T.cell_id[l] = c;
  atomicAdd(&T.counts[c], 1);
}

// starts = exclusive scan of counts (3 passes, as the weights)
__global__ void __launch_bounds__(kScanBlock) plan_cells_partial(PlanTree t0, PlanTree t1, int ncells) {
  __shared__ int lds[kMaxWaves + 1];
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  const int i0 = blockIdx.x * kPlanScanItems + threadIdx.x * kScanPer;
  int v[kScanPer];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = s;
    s += (i0 + k < ncells) ? T.counts[i0 + k] : 0;
  }
  int tot;
  const int ex = block_excl_scan(s, &tot, lds);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (i0 + k < ncells) T.starts[i0 + k] = ex + v[k];
  if (threadIdx.x == 0) T.cbsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanBlock) plan_cells_top(PlanTree t0, PlanTree t1, int nb) {
  scan_block_sums<int>(tree_of(t0, t1, blockIdx.x).cbsum, nb);
}

__global__ void __launch_bounds__(kScanBlock) plan_cells_add(PlanTree t0, PlanTree t1, int ncells) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  const int add = T.cbsum[blockIdx.x];
  const int base = blockIdx.x * kPlanScanItems;
  for (int i = base + threadIdx.x; i < base + kPlanScanItems && i < ncells; i += kScanBlock) {
    const int s = T.starts[i] + add;
    T.starts[i] = s;
    T.cursor[i] = s;
  }
}

__global__ void __launch_bounds__(256) plan_grid_scatter(int ndof, PlanTree t0, PlanTree t1, int cap) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= *T.n) return;
  const int p = atomicAdd(&T.cursor[T.cell_id[l]], 1);
  for (int i = 0; i < ndof; ++i) T.sorted[p + (size_t)cap * i] = T.q[l + (size_t)cap * i];
  T.sid[p] = l;
}

// ---- shortest paths from the roots (replaces the host's do_dijkstra) ----------
// Bellman-Ford over each tree's edges: distances as float bits (non-negative
// floats order as their bits), lowered by atomicMin; relaxation rounds until
// one changes nothing.  Then every node's parent is the lowest-numbered
// neighbour u with dist[u] + w == dist[v] and dist[u] < dist[v], and the path
// is followed from the connection node back to the root.
constexpr int kInfBits = 0x7f800000;

__global__ void __launch_bounds__(256) plan_sssp_init(PlanTree t0, PlanTree t1, int cap, int* __restrict__ dist,
                                                      int* __restrict__ parent) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= *T.n) return;
  dist[(size_t)blockIdx.y * cap + l] = l == 0 ? 0 : kInfBits;
  parent[(size_t)blockIdx.y * cap + l] = 0x7fffffff;
}

__device__ __forceinline__ int tree_edges(const PlanTree& T) {
  const unsigned c = *T.ecount;
  return (int)(c < (unsigned)T.ecap ? c : (unsigned)T.ecap);
}

__global__ void __launch_bounds__(256) plan_sssp_relax(PlanTree t0, PlanTree t1, int cap, int* __restrict__ dist,
                                                       int* __restrict__ changed) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  int* D = dist + (size_t)blockIdx.y * cap;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (x < tree_edges(T)) {
    const int u = T.efrom[x], v = T.eto[x];
    const float w = T.ew[x];
    const float du = __int_as_float(D[u]), dv = __int_as_float(D[v]);
    const float a = du + w, c = dv + w;
    if (a < dv) {
      atomicMin(&D[v], __float_as_int(a));
      ch = true;
    }
    if (c < du) {
      atomicMin(&D[u], __float_as_int(c));
      ch = true;
    }
  }
  if (__ballot(ch) && (threadIdx.x & 63) == 0) changed[blockIdx.y] = 1;
}

__global__ void __launch_bounds__(256) plan_sssp_parent(PlanTree t0, PlanTree t1, int cap, const int* __restrict__ dist,
                                                        int* __restrict__ parent) {
  const PlanTree& T = tree_of(t0, t1, blockIdx.y);
  const int* D = dist + (size_t)blockIdx.y * cap;
  int* P = parent + (size_t)blockIdx.y * cap;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= tree_edges(T)) return;
  const int u = T.efrom[x], v = T.eto[x];
  const float w = T.ew[x];
  const float du = __int_as_float(D[u]), dv = __int_as_float(D[v]);
  if (du < dv && du + w == dv) atomicMin(&P[v], u);
  if (dv < du && dv + w == du) atomicMin(&P[u], v);
}

// One thread per tree: the chain from `from` (conn_left, conn_right) back to
// node 0, at most maxlen nodes; len[tree] = -1 when a parent is missing.
__global__ void plan_path_chase(int cap, const int* __restrict__ parent, const PlanState* __restrict__ state,
                                int maxlen, int* __restrict__ chain, int* __restrict__ len) {
  const int g = threadIdx.x;
  if (g >= 2) return;
  const int* P = parent + (size_t)g * cap;
  int v = g == 0 ? state->conn_left : state->conn_right;
  int n = 0;
  for (;;) {
    if (n >= maxlen || v < 0 || v >= cap) {
      n = -1;
      break;
    }
    chain[(size_t)g * maxlen + n++] = v;
    if (v == 0) break;
    v = P[v];
  }
  len[g] = n;
}

// Coordinates of the chain nodes of both trees, AoS: out[(g*maxlen + i)*ndof + c].
__global__ void __launch_bounds__(256) plan_path_gather(int ndof, PlanTree t0, PlanTree t1, int cap, int maxlen,
                                                        const int* __restrict__ chain, const int* __restrict__ len,
                                                        float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  if (i >= len[g]) return;
  const PlanTree& T = tree_of(t0, t1, g);
  const int v = chain[(size_t)g * maxlen + i];
  for (int c = 0; c < ndof; ++c) out[((size_t)g * maxlen + i) * ndof + c] = T.q[v + (size_t)cap * c];
}

inline unsigned blocks_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

hipError_t launch_plan_weights(PlanTree t0, PlanTree t1, int nb, hipStream_t st) {
  hipLaunchKernelGGL(plan_weights_partial, dim3(nb, 2), dim3(kScanBlock), 0, st, t0, t1);
  hipLaunchKernelGGL(plan_weights_top, dim3(2), dim3(kScanBlock), 0, st, t0, t1, nb);
  hipLaunchKernelGGL(plan_weights_add, dim3(nb, 2), dim3(kScanBlock), 0, st, t0, t1);
  return hipGetLastError();
}

hipError_t launch_plan_sample(int ndof, int batch, PlanTree t0, PlanTree t1, int cap, int nb, float D,
                              unsigned long long key, float* cand, hipStream_t st) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(plan_sample_kernel, dim3(blocks_for(batch, 256)), dim3(256), 0, st, ndof, batch, t0, t1, cap, nb,
                     D, key, cand);
  return hipGetLastError();
}

hipError_t launch_plan_bias_mask(const float* prob, int batch, unsigned char* skip, hipStream_t st) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(plan_bias_mask_kernel, dim3(blocks_for(batch, 256)), dim3(256), 0, st, prob, batch, skip);
  return hipGetLastError();
}

hipError_t launch_plan_select(int ndof, int batch, const int* codes, const unsigned char* skip, const float* cand,
                              int want0, int want1, int num, float* qnew, int* tile, StepInfo* info, PlanState* state,
                              hipStream_t st) {
  const int ntiles = (int)blocks_for(batch, kScanBlock);
  hipLaunchKernelGGL(plan_select_count, dim3(ntiles), dim3(kScanBlock), 0, st, batch, codes, skip, tile);
  hipLaunchKernelGGL(plan_select_prefix, dim3(1), dim3(kScanBlock), 0, st, batch, ntiles, want0, want1, tile, info,
                     state);
  hipLaunchKernelGGL(plan_select_write, dim3(ntiles), dim3(kScanBlock), 0, st, ndof, batch, codes, skip, cand, want0,
                     want1, num, tile, qnew);
  return hipGetLastError();
}

hipError_t launch_plan_near(int ndof, int num, const float* qnew, StepInfo* info, PlanTree t0, PlanTree t1, int cap,
                            const CellGrid& g, float D2, int* cnt, int* off, int* sc_off, int nbuf, int budget,
                            int* enode, float* ed2, int* egroup, hipStream_t st) {
  const int groups = 2 * num;
  const int keep = std::max(1, std::min(kGroupKeep, nbuf / groups));
  // info->E_all was reset by the step's selection
  if (budget <= 0)
    hipLaunchKernelGGL(plan_near_kernel<kNearAppend>, dim3(groups), dim3(256), 0, st, ndof, num, qnew, info, t0, t1,
                       cap, g, D2, cnt, off, nbuf, keep, enode, ed2, egroup);
  else
    hipLaunchKernelGGL(plan_near_kernel<kNearGroup>, dim3(groups), dim3(256), 0, st, ndof, num, qnew, info, t0, t1,
                       cap, g, D2, cnt, off, nbuf, keep, enode, ed2, egroup);
  hipLaunchKernelGGL(plan_near_done_kernel, dim3(1), dim3(kScanBlock), 0, st, groups, cnt, nbuf, budget, sc_off,
                     info);
  return hipGetLastError();
}

hipError_t launch_plan_paths_init(PlanTree t0, PlanTree t1, int cap, int* dist, int* parent, hipStream_t st) {
  hipLaunchKernelGGL(plan_sssp_init, dim3(blocks_for(cap, 256), 2), dim3(256), 0, st, t0, t1, cap, dist, parent);
  return hipGetLastError();
}

hipError_t launch_plan_paths_relax(PlanTree t0, PlanTree t1, int cap, int max_edges, int* dist, int* changed,
                                   hipStream_t st) {
  if (max_edges <= 0) return hipSuccess;
  hipLaunchKernelGGL(plan_sssp_relax, dim3(blocks_for(max_edges, 256), 2), dim3(256), 0, st, t0, t1, cap, dist,
                     changed);
  return hipGetLastError();
}

hipError_t launch_plan_paths_finish(int ndof, PlanTree t0, PlanTree t1, int cap, int max_edges, const int* dist,
                                    int* parent, const PlanState* state, int maxlen, int* chain, int* len, float* coords,
                                    hipStream_t st) {
  if (max_edges > 0)
    hipLaunchKernelGGL(plan_sssp_parent, dim3(blocks_for(max_edges, 256), 2), dim3(256), 0, st, t0, t1, cap, dist,
                       parent);
  hipLaunchKernelGGL(plan_path_chase, dim3(1), dim3(64), 0, st, cap, parent, state, maxlen, chain, len);
  hipLaunchKernelGGL(plan_path_gather, dim3(blocks_for(maxlen, 256), 2), dim3(256), 0, st, ndof, t0, t1, cap, maxlen,
                     chain, len, coords);
  return hipGetLastError();
}

hipError_t launch_plan_score_points(int ndof, int E, int nbuf, int budget, int S8, const int* cnt, const int* off,
                                    const int* sc_off, const int* egroup, const int* enode, const float* qnew, int num,
                                    PlanTree t0, PlanTree t1, int cap, float* out, long long ns, hipStream_t st) {
  if (E <= 0) return hipSuccess;
  hipLaunchKernelGGL(plan_score_points_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, st, ndof, E, nbuf, budget, S8,
                     cnt, off, sc_off, egroup, enode, qnew, num, t0, t1, cap, out, ns);
  return hipGetLastError();
}

hipError_t launch_plan_rank(int E, int nbuf, int budget, int S8, const int* cnt, const int* off, const int* sc_off,
                            const int* egroup, const float* ed2, const float* prob, float* score,
                            unsigned char* skip, PlanState* state, hipStream_t st) {
  if (E <= 0) return hipSuccess;
  hipLaunchKernelGGL(plan_edge_score_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, st, E, nbuf, budget, S8, cnt, off,
                     sc_off, egroup, prob, score);
  hipLaunchKernelGGL(plan_rank_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, st, E, nbuf, budget, cnt, off, egroup,
                     ed2, score, skip, state);
  return hipGetLastError();
}

hipError_t launch_plan_gather(int ndof, int E, const int* egroup, const int* enode, const float* qnew, int num,
                              PlanTree t0, PlanTree t1, int cap, float* qs, float* qe, hipStream_t st) {
  if (E <= 0) return hipSuccess;
  hipLaunchKernelGGL(plan_gather_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, st, ndof, E, egroup, enode, qnew, num,
                     t0, t1, cap, qs, qe);
  return hipGetLastError();
}

hipError_t launch_plan_edges_done(int E, const int* egroup, const int* res, const unsigned char* skip, int* free_cnt,
                                  const unsigned long long* total, const int* max_t, int maxT_bound,
                                  PlanState* state, hipStream_t st) {
  if (E <= 0) return hipSuccess;
  hipLaunchKernelGGL(plan_edges_done_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, st, E, egroup, res, skip,
                     free_cnt, total, max_t, maxT_bound, state);
  return hipGetLastError();
}

hipError_t launch_plan_insert(int ndof, int num, int E, const StepInfo* info, const int* free_cnt, int* pos,
                              const float* qnew, const int* egroup, const int* enode, const float* ed2, const int* res,
                              PlanTree t0, PlanTree t1, int cap, PlanState* state, hipStream_t st) {
  hipLaunchKernelGGL(plan_insert_plan_kernel, dim3(1), dim3(kScanBlock), 0, st, num, info, free_cnt, pos, cap, state);
  hipLaunchKernelGGL(plan_insert_nodes_kernel, dim3(blocks_for(2LL * num, 256)), dim3(256), 0, st, ndof, num, pos,
                     free_cnt, qnew, t0, t1, cap);
  if (E > 0)
    hipLaunchKernelGGL(plan_insert_edges_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, st, E, egroup, enode, ed2, res,
                       pos, t0, t1, state);
  return hipGetLastError();
}

hipError_t launch_plan_index(int ndof, PlanTree t0, PlanTree t1, int cap, const CellGrid& g, hipStream_t st) {
  const int ncells = g.ncells;
  const int nbc = plan_scan_blocks(ncells);
  unsigned fb = blocks_for(ncells, 256);
  if (fb > 1024) fb = 1024;
  hipLaunchKernelGGL(plan_fill_counts, dim3(fb, 2), dim3(256), 0, st, t0, t1, ncells);
  hipLaunchKernelGGL(plan_grid_count, dim3(blocks_for(cap, 256), 2), dim3(256), 0, st, t0, t1, cap, g);
  hipLaunchKernelGGL(plan_cells_partial, dim3(nbc, 2), dim3(kScanBlock), 0, st, t0, t1, ncells);
  hipLaunchKernelGGL(plan_cells_top, dim3(2), dim3(kScanBlock), 0, st, t0, t1, nbc);
  hipLaunchKernelGGL(plan_cells_add, dim3(nbc, 2), dim3(kScanBlock), 0, st, t0, t1, ncells);
  hipLaunchKernelGGL(plan_grid_scatter, dim3(blocks_for(cap, 256), 2), dim3(256), 0, st, ndof, t0, t1, cap);
  return hipGetLastError();
}

}  // namespace gprm
