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

// One block: slot order, at most want[g] free candidates of tree g.
__global__ void __launch_bounds__(kScanBlock)
    plan_select_kernel(int ndof, int batch, const int* __restrict__ codes, const unsigned char* __restrict__ skip,
                       const float* __restrict__ cand, int want0, int want1, int num, float* __restrict__ qnew,
                       StepInfo* info, PlanState* state) {
  __shared__ int lds[kMaxWaves + 1];
  const int per = (batch + kScanBlock - 1) / kScanBlock;
  const int k0 = threadIdx.x * per, k1 = min(batch, k0 + per);
  int a0 = 0, a1 = 0, dr = 0;
  for (int k = k0; k < k1; ++k) {
    const bool sk = skip && skip[k];
    dr += sk;
    if (!sk && codes[k] == 0) ((k & 1) ? a1 : a0) += 1;
  }
  int A0, A1, DR;
  const int r0s = block_excl_scan(a0, &A0, lds);
  const int r1s = block_excl_scan(a1, &A1, lds);
  (void)block_excl_scan(dr, &DR, lds);
  // kept candidates of this thread, then their order among all kept
  int r0 = r0s, r1 = r1s, kept = 0;
  for (int k = k0; k < k1; ++k) {
    if ((skip && skip[k]) || codes[k] != 0) continue;
    if (k & 1) kept += r1++ < want1;
    else kept += r0++ < want0;
  }
  int nnew;
  int j = block_excl_scan(kept, &nnew, lds);
  r0 = r0s;
  r1 = r1s;
  for (int k = k0; k < k1; ++k) {
    if ((skip && skip[k]) || codes[k] != 0) continue;
    const bool take = (k & 1) ? r1++ < want1 : r0++ < want0;
    if (!take) continue;
    for (int i = 0; i < ndof; ++i) qnew[j + (size_t)num * i] = cand[k + (size_t)batch * i];
    ++j;
  }
  if (threadIdx.x == 0) {
    info->nnew = nnew;
    info->accepted = A0 + A1;
    atomicAdd(&state->point_checks, (unsigned long long)(batch - DR));
    atomicAdd(&state->dropped, (unsigned long long)DR);
  }
}

// ---- radius search -----------------------------------------------------------
__device__ __forceinline__ int grid_coord(const CellGrid& g, int d, float x) {
  int c = (int)floorf((x - g.lo[d]) * g.inv[d]);
  return c < 0 ? 0 : (c >= g.cells[d] ? g.cells[d] - 1 : c);
}

constexpr int pow3(int d) { return d == 0 ? 1 : 3 * pow3(d - 1); }
constexpr int kNeighbourCells = pow3(kGridDims);
static_assert(kNeighbourCells <= 256, "one thread per neighbour cell in a 256-thread block");

template <bool EMIT>
__global__ void __launch_bounds__(256)
    plan_near_kernel(int ndof, int num, const float* __restrict__ qnew, const StepInfo* __restrict__ info,
                     PlanTree t0, PlanTree t1, int cap, CellGrid g, float D2, int* __restrict__ cnt,
                     const int* __restrict__ off, int nbuf, int* __restrict__ enode, float* __restrict__ ed2,
                     int* __restrict__ egroup) {
  constexpr int NK = kNeighbourCells;
  __shared__ int s_start[NK], s_pre[NK + 1];
  __shared__ int s_wave[4];
  const int b = blockIdx.x, j = b >> 1;
  if (j >= info->nnew) {
    if (!EMIT && threadIdx.x == 0) cnt[b] = 0;
    return;
  }
  const PlanTree& T = tree_of(t0, t1, b & 1);
  float qr[kMaxDofDev];
  for (int i = 0; i < ndof; ++i) qr[i] = qnew[j + (size_t)num * i];
  if (threadIdx.x < NK) {
    // neighbour k: offset digit(d) - 1 in grid dimension d; a dimension
    // without grid (d >= dims) takes only offset 0
    int rem = threadIdx.x, c = 0;
    bool ok = true;
    int digit[kGridDims];
    for (int d = 0; d < kGridDims; ++d) {
      digit[d] = rem % 3;
      rem /= 3;
    }
    for (int d = kGridDims - 1; d >= 0; --d) {
      if (d >= g.dims) {
        if (digit[d] != 1) ok = false;
        continue;
      }
      const int x = grid_coord(g, d, qr[d]) + digit[d] - 1;
      if (x < 0 || x >= g.cells[d]) ok = false;
      c = c * g.cells[d] + x;
    }
    s_start[threadIdx.x] = ok ? T.starts[c] : 0;
    s_pre[threadIdx.x + 1] = ok ? T.counts[c] : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_pre[0] = 0;
    for (int k = 0; k < NK; ++k) s_pre[k + 1] += s_pre[k];
  }
  __syncthreads();
  const int total = s_pre[NK];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int mine = 0;                            // COUNT: this thread's hits
  const int gbase = EMIT ? off[b] : 0;     // EMIT: the group's first slot
  int run = 0;                             // EMIT: hits of the earlier iterations
  for (int base = 0; base < total; base += blockDim.x) {
    const int idx = base + threadIdx.x;
    float normsq = 0.0f;
    bool hit = false;
    int node = 0;
    if (idx < total) {
      int k = 0;  // the neighbour cell holding idx: last k with s_pre[k] <= idx
      for (int step = 128; step > 0; step >>= 1)
        if (k + step < NK && s_pre[k + step] <= idx) k += step;
      const int pos = s_start[k] + (idx - s_pre[k]);
      for (int i = 0; i < ndof; ++i) {
        const float diff = T.sorted[pos + (size_t)cap * i] - qr[i];
        normsq += diff * diff;
      }
      hit = normsq < D2;
      node = T.sid[pos];
    }
    if (!EMIT) {
      mine += hit;
      continue;
    }
    const unsigned long long m = __ballot(hit);
    if (lane == 0) s_wave[w] = __popcll(m);
    __syncthreads();
    int before = 0, all = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      before += i < w ? s_wave[i] : 0;
      all += s_wave[i];
    }
    if (hit) {
      const int o = gbase + run + before + __popcll(m & ((1ull << lane) - 1ull));
      if (o < nbuf) {
        enode[o] = node;
        ed2[o] = normsq;
        egroup[o] = b;
      }
    }
    run += all;
    __syncthreads();  // s_wave is rewritten by the next iteration
  }
  if (!EMIT) {
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    for (int d = 32; d >= 1; d >>= 1) mine += __shfl_xor(mine, d, 64);
    if (lane == 0) atomicAdd(&s_cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) cnt[b] = s_cnt;
  }
}

// Edges of group b actually stored (the first nbuf edges of the step are).
__device__ __forceinline__ int stored_edges(const int* off, const int* cnt, int b, int nbuf) {
  const int a = min(off[b], nbuf), e = min(off[b] + cnt[b], nbuf);
  return e - a;
}

__global__ void __launch_bounds__(kScanBlock)
    plan_scan_groups_kernel(int groups, const int* __restrict__ cnt, int nbuf, int budget, int* __restrict__ off,
                            int* __restrict__ sc_off, StepInfo* info) {
  __shared__ int lds[kMaxWaves + 1];
  const int per = (groups + kScanBlock - 1) / kScanBlock;
  const int b0 = threadIdx.x * per, b1 = min(groups, b0 + per);
  int s = 0;
  for (int b = b0; b < b1; ++b) s += cnt[b];
  int all;
  int run = block_excl_scan(s, &all, lds);
  for (int b = b0; b < b1; ++b) {
    off[b] = run;
    run += cnt[b];
  }
  if (threadIdx.x == 0) off[groups] = all;
  __syncthreads();
  int S = 0;
  if (budget > 0) {
    s = 0;
    for (int b = b0; b < b1; ++b) {
      const int w = stored_edges(off, cnt, b, nbuf);
      s += w > budget ? w : 0;
    }
    run = block_excl_scan(s, &S, lds);
    for (int b = b0; b < b1; ++b) {
      sc_off[b] = run;
      const int w = stored_edges(off, cnt, b, nbuf);
      run += w > budget ? w : 0;
    }
    if (threadIdx.x == 0) sc_off[groups] = S;
  }
  if (threadIdx.x == 0) {
    info->E_all = all;
    info->E = all < nbuf ? all : nbuf;
    info->S = S;
  }
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
      const int o = off[b];
      const float se = score[e], de = ed2[e];
      int rank = 0;
      for (int x = o; x < o + w; ++x) {
        const float sx = score[x], dx = ed2[x];
        rank += (sx < se) || (sx == se && (dx < de || (dx == de && x < e)));
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
  if (e >= E || res[e] != 0) return;
  const int b = egroup[e];
  const int p = pos[b];
  if (p < 0) return;
  const PlanTree& T = tree_of(t0, t1, b & 1);
  const int node = enode[e];
  atomicAdd(&T.surr[node], 1);
  const unsigned x = atomicAdd(T.ecount, 1u);
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
                              int want0, int want1, int num, float* qnew, StepInfo* info, PlanState* state,
                              hipStream_t st) {
  hipLaunchKernelGGL(plan_select_kernel, dim3(1), dim3(kScanBlock), 0, st, ndof, batch, codes, skip, cand, want0, want1,
                     num, qnew, info, state);
  return hipGetLastError();
}

hipError_t launch_plan_near(bool emit, int ndof, int num, const float* qnew, const StepInfo* info, PlanTree t0,
                            PlanTree t1, int cap, const CellGrid& g, float D2, int* cnt, const int* off, int nbuf,
                            int* enode, float* ed2, int* egroup, hipStream_t st) {
  if (num <= 0) return hipSuccess;
  if (emit)
    hipLaunchKernelGGL(plan_near_kernel<true>, dim3(2 * num), dim3(256), 0, st, ndof, num, qnew, info, t0, t1, cap, g,
                       D2, cnt, off, nbuf, enode, ed2, egroup);
  else
    hipLaunchKernelGGL(plan_near_kernel<false>, dim3(2 * num), dim3(256), 0, st, ndof, num, qnew, info, t0, t1, cap, g,
                       D2, cnt, off, nbuf, enode, ed2, egroup);
  return hipGetLastError();
}

hipError_t launch_plan_scan_groups(int groups, const int* cnt, int nbuf, int budget, int* off, int* sc_off,
                                   StepInfo* info, hipStream_t st) {
  hipLaunchKernelGGL(plan_scan_groups_kernel, dim3(1), dim3(kScanBlock), 0, st, groups, cnt, nbuf, budget, off, sc_off,
                     info);
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
