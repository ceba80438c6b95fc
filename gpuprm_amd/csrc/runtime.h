// runtime.h -- internal interface between the host runtime (gprm_api.cpp) and
// the HIP kernels (kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gprm {

// Threads per block of the kinematics kernels.  An edge launch's kinematics
// runs beside the previous launch's separating-vector kernel; with 1024
// rather than 256 threads per block the edge benchmark measured 4.93 against
// 4.98 ms per step.
constexpr int kFkBlock = 1024;
constexpr int kMaxIter = 10;        // collision4.h:17-18 (max_for_loop, max_for_loop_whole)
constexpr int kMaxDofDev = 16;

// Robot parameters (modified DH, robot.h:10-22) and joint bounds.
struct DevRobot {
  int ndof;
  float a[kMaxDofDev], alpha[kMaxDofDev], q[kMaxDofDev], d[kMaxDofDev];
  float mins[kMaxDofDev], maxs[kMaxDofDev];
  int types[kMaxDofDev];
  // cos/sin of the constant twists alpha and of 0 (for T0 = set(0,0,0,0)),
  // evaluated once on the host by the same restated glibc routines the
  // kernels use, so the bits are those the kernels would compute
  float ca[kMaxDofDev], sa[kMaxDofDev];
  float c0, s0;
};

// One contiguous device image of a scene, staged whole into LDS by the
// separating-vector kernel.  Sections (byte offsets, 16-byte aligned):
//   verts    float4[sumn]  vertex (x,y,z) in the polytope's own frame; w holds
//                          the bits of (adjacency row offset << 8) | degree
//   pairs    int4[M]       (vbase P | vbase Q << 16, frame P | frame Q << 8,
//                          CRS index, polytope P | polytope Q << 16), grouped
//                          by the frames they need
//   adj      uint16[...]   neighbour vertex ids (scene-global), rows as stored,
//                          each padded to a multiple of 4 with the row's vertex
//   spheres  float4[2N]    per polytope, own frame: (centre, sphere radius),
//                          (box half extents, |centre|_1 + radius), rounded
//                          outwards
struct SceneImg {
  const int4* blob;        // device pointer to the image
  int bytes;               // image size, multiple of 16
  int off_verts, off_pairs, off_adj, off_spheres;
  int sumn, summ, M, N;
};

// Per-launch arguments of the separating-vector kernel.
struct SvaArgs {
  SceneImg img;
  int ndof;
  const float* frames;     // link frames, link-major and quad-planar: float4 j (0..2) of
                           // link s (1..ndof) of item i at frames4[((s-1) 3 + j) fstride + i]
  long long fstride;       // item stride of frames (n of the kinematics launch, or its cap)
  float4* hist;            // [kMaxIter][grid*block] per-lane iteration history
  int nitems;              // item count, or -1: read *nitems_dev
  const int* nitems_dev;
  unsigned int* counter;   // work counter (zero before launch)
  int pairs_per_unit;      // pairs of one work unit (a wave claims one chunk of them
                           // for 64 consecutive items)
  int units_per_item;      // ceil(M / pairs_per_unit)
  // point mode
  int* word;               // [item] first colliding pair and code, see fk_points_kernel
  // point mode, optional (both null: a separate finalize launch writes the
  // codes): the wave that completes a group's last unit writes the group's
  // codes from the words.  group_done: [groups of 64 items] completed units
  // (zero before launch).
  int* group_done;
  int* codes;
  // edge mode
  const int* item_edge;    // [item] owning edge
  int* edge_res;           // [edge] nonzero once any sample collides
  // optional counters (may be null): kTallySlots slots of kTallyCount, block
  // b adds its sums into slot b mod kTallySlots, spreading the atomics; a
  // slot holds [0] pair tests ending separated, [1] pair tests ending in a
  // collision, [2] pair tests skipped by early exit, [3] units completed (per
  // lane), [4] iterations run, [5] pair tests decided separated by the bound
  // test
  unsigned long long* tally;
  // single-scene launches: wave w of the grid takes unit w without a claim,
  // and claims from the counter start after the grid's waves
  int static_first;
};
constexpr int kTallyCount = 6;
constexpr int kTallySlots = 64;

// One scene of a multi-scene separating-vector launch (device memory).
struct SceneTask {
  SvaArgs a;                    // the scene's arguments; hist is shared by all tasks
  unsigned* done;               // units finished (zero before launch)
  unsigned long long* t_done;   // device clock (wall_clock64) when its last unit finished
};

// One scene of a multi-scene kinematics / finalize launch (device memory).
// Plane stride of the quad-planar frames for n items: n rounded up to 64,
// so every plane starts 1 KB-aligned and a wave's float4 store (64 lanes,
// 1 KB) covers whole 128-byte lines.  With an odd stride the edge benchmark's
// third kinematics launch ran 20% longer.
inline long long frame_stride(long long n) { return (n + 63) / 64 * 64; }

struct FkTask {
  DevRobot rb;
  const float* q;               // SoA, component j of item i at q[i + offset*j]
  long long offset;
  int n;
  int block0;                   // first block of this task (tasks sorted by block0)
  long long fstride;            // frame plane stride, frame_stride(n)
  float* frames;                // link-major, item stride fstride
  int* word;                    // [n] result words
  int* codes;                   // [n] codes written by the finalize kernel
};

// Cell index of the radius search: nodes binned by their first `dims`
// coordinates (cells at least D wide, at most 64 per dimension; up to
// kGridDims dimensions while the cells number at most kGridMaxCells).  A
// query visits the 3^dims cells around its own.  On the 7-DOF planner task
// (D = 0.5: 12 cells per dimension) five binned coordinates leave about a
// quarter of the candidates of four.
constexpr int kGridDims = 5;
constexpr int kGridMaxCells = 1 << 20;
struct CellGrid {
  int dims;
  int cells[kGridDims];
  float lo[kGridDims], inv[kGridDims];   // cell = clamp(floor((x - lo) * inv), 0, cells - 1)
  int ncells;
};

// Cell coordinate of x in grid dimension d (clamped to the grid).
__device__ __forceinline__ int grid_coord(const CellGrid& g, int d, float x) {
  int c = (int)floorf((x - g.lo[d]) * g.inv[d]);
  return c < 0 ? 0 : (c >= g.cells[d] ? g.cells[d] - 1 : c);
}

// Squared lower bound, in dimension d, on the distance from x to any node of
// the neighbour cell at offset digit - 1 from x's own cell: the gap to the
// face the two cells share (nodes of the cell below have (y - lo) inv < c,
// of the cell above >= c + 1, as grid_coord bins them).  The gap is first
// shortened by a thousandth of a cell, far above the rounding of these
// quantities, so a cell whose bounds sum to D^2 or more holds no node the
// exact test (distance^2 < D^2) would match.
__device__ __forceinline__ float cell_gap2(const CellGrid& g, int d, float x, int digit) {
  if (digit == 1) return 0.0f;
  const float t = (x - g.lo[d]) * g.inv[d];
  const float c = (float)grid_coord(g, d, x);
  const float gap = (digit == 0 ? t - c : c + 1.0f - t) - 1e-3f;
  if (!(gap > 0.0f)) return 0.0f;
  const float len = gap / g.inv[d];
  return len * len;
}

// Squared distances out[u] from q to stored nodes pos[u], u < P (coordinate
// i of node p at coords[p + stride i]), each summed in coordinate order like
// the reference's float loop.  The first eight coordinates of all P nodes
// are loaded before any is used, so their loads are in flight together (a
// loop over a run-time ndof would wait for each load in turn).  Every pos[u]
// must be a stored node.
template <int P>
__device__ __forceinline__ void node_dist2(const float* __restrict__ coords, size_t stride, const int* pos, int ndof,
                                           const float* q, float* out) {
  constexpr int kFirst = 8;
  float x[P][kFirst];
#pragma unroll
  for (int i = 0; i < kFirst; ++i)
#pragma unroll
    for (int u = 0; u < P; ++u) x[u][i] = i < ndof ? coords[pos[u] + stride * i] : 0.0f;
#pragma unroll
  for (int u = 0; u < P; ++u) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < kFirst; ++i)
      if (i < ndof) {
        const float diff = x[u][i] - q[i];
        s += diff * diff;
      }
    for (int i = kFirst; i < ndof; ++i) {
      const float diff = coords[pos[u] + stride * i] - q[i];
      s += diff * diff;
    }
    out[u] = s;
  }
}

// Result of an edge left unchecked by EdgeArgs::skip (nonzero: never taken
// as free).
constexpr int kEdgeSkipped = 0x40000000;

// Edge work description.
struct EdgeArgs {
  const float* qs; const float* qe; long long offset;  // SoA, component i of edge k at [k + offset*i]
  int nedges;
  // optional [edge]: nonzero = leave the edge unchecked (testnum 0, result
  // kEdgeSkipped; the planner's edge budget)
  const unsigned char* skip;
  const int* testnum;      // [edge]
  const int* stride;       // [edge] sample order: position j -> (j*stride) mod testnum
  int* edge_res;           // [edge]
};

// One scene's part of a multi-scene edge call (device memory).  Prologue
// tasks (one per scene) cover edges [0, ea.nedges) and write testnum,
// stride, the scene's largest testnum and its sample total; launch-step
// tasks cover edges [e0, e1) and order positions [lo, hi): the compaction
// appends their items (item_edge, item_pos, *nitems, at most cap) and the
// kinematics writes their frames (link-major, item stride cap).
struct EdgeTask {
  EdgeArgs ea;
  DevRobot rb;
  float dq;
  int* testnum;
  int* stride;
  int* max_testnum;
  unsigned long long* total;
  int e0, e1, lo, hi;
  int* item_edge;
  int* item_pos;
  int* nitems;
  int cap;                      // item capacity, a multiple of 64 (also the frame plane stride)
  float* frames;
  int block0;    // first block of the task in the prologue / compaction grid (tasks sorted)
  int fblock0;   // first block of the task in the kinematics grid (tasks sorted)
};
hipError_t launch_edge_prologue_multi(const EdgeTask* d_tasks, int ntasks, int nblocks, hipStream_t st);
hipError_t launch_edge_compact_multi(const EdgeTask* d_tasks, int ntasks, int nblocks, hipStream_t st);
hipError_t launch_edge_fk_multi(const EdgeTask* d_tasks, int ntasks, int nblocks, int ndof, hipStream_t st);

// Deferral rings of the separating-vector kernel: per wave, a ring of
// sva_ring_cap(block) entries of 12 bytes (item, pair, edge).  Larger rings
// start the full tests in larger batches, which shortens their tails; the
// 1024-thread variant keeps smaller rings to leave LDS to large scenes.
constexpr int sva_ring_cap(int block) { return block >= 1024 ? 128 : 256; }
constexpr size_t sva_ring_bytes(int block) { return (size_t)(block / 64) * (size_t)sva_ring_cap(block) * 12; }
// Dynamic LDS of a separating-vector block of `block` threads on this scene:
// its rings and, unless the kernel reads the scene from the global image
// (gscene), a copy of the image.
size_t sva_lds_bytes(const SceneImg& img, int block, int gscene = 0);
int sva_max_lds_bytes();

// Kinematics of n configurations; frames link-major with item stride n.
// zero_counter (may be null) is set to zero by the launch (the work counter
// of the separating-vector launch that follows).  skip (may be null): items
// with skip[i] != 0 get the out-of-bounds word, so the separating-vector
// kernel tests none of their pairs (code 1).
hipError_t launch_fk_points(const DevRobot& rb, const float* q, long long offset, int n, float* frames,
                            long long fstride, int* word, int check_bounds, unsigned* zero_counter, hipStream_t st,
                            const unsigned char* skip = nullptr, int* group_done = nullptr);
hipError_t launch_finalize_points(const int* word, int* codes, int n, hipStream_t st);
// The stored link frames of n items (and the base frame) unpacked to
// out[(i (ndof + 1) + s) 12 + 3 column + row].
hipError_t launch_frames_unpack(const DevRobot& rb, const float* frames, long long fstride, int n, float* out,
                               hipStream_t st);
struct SvaShape;
hipError_t launch_sva(const SvaArgs& a, int edge_mode, const SvaShape& sh, hipStream_t st);
// Multi-scene launches (d_tasks in device memory).  The kinematics and
// finalize grids have nblocks = sum over tasks of ceil(n / kFkBlock) blocks.
hipError_t launch_fk_points_multi(const FkTask* d_tasks, int ntasks, int nblocks, int ndof, hipStream_t st);
hipError_t launch_finalize_points_multi(const FkTask* d_tasks, int ntasks, int nblocks, hipStream_t st);
hipError_t launch_sva_multi(const SceneTask* d_tasks, int ntasks, int edge_mode, const SvaShape& sh, size_t img_bytes,
                            unsigned long long* t_start, unsigned long long* tally, hipStream_t st);
hipError_t launch_edge_prologue(const DevRobot& rb, float dq, EdgeArgs ea, int* testnum_out,
                                int* stride_out, int* max_testnum, unsigned long long* total,
                                hipStream_t st);
hipError_t launch_edge_compact(EdgeArgs ea, int e0, int e1, int lo, int hi, int* item_edge,
                               int* item_pos, int* nitems, hipStream_t st);
// Kinematics of the *nitems (<= cap) compacted samples; frames link-major
// with item stride cap.
hipError_t launch_edge_fk(const DevRobot& rb, EdgeArgs ea, const int* item_edge, const int* item_pos,
                          const int* nitems, int cap, float* frames, hipStream_t st);  // cap: plane stride
// Small point-check calls: up to kSmallMax configurations passed by value,
// one block each, codes written to out (device-visible, e.g. pinned host
// memory).  point_small_prepare sets the kernels' LDS limit once per scene.
constexpr int kSmallMax = 8;
struct SmallPoints {
  int n;
  float q[kSmallMax][kMaxDofDev];
};
hipError_t point_small_prepare(const SceneImg& img);
hipError_t launch_point_small(const SceneImg& img, const DevRobot& rb, const SmallPoints& qs, int* out,
                              hipStream_t st);
hipError_t launch_fill_i32(int* p, int v, long long n, hipStream_t st);
CellGrid make_cell_grid(int ndof, const float* mins, const float* maxs, float D);
// Counting sort of n nodes (SoA, stride noff) into the cell index: counts and
// starts per cell (g.ncells), cursor scratch (g.ncells), cell_id scratch (n),
// sorted coordinates (SoA, stride n) and node ids (n).
hipError_t launch_grid_build(const CellGrid& g, int ndof, const float* nodes, int n, long long noff, int* cell_id,
                             int* counts, int* starts, int* cursor, float* sorted, int* sid, hipStream_t st);
// Radius neighbours over the cell index: every (query j, node l) with
// |node_l - query_j|^2 < D2, appended to (out_q, out_node, out_d2) up to cap
// entries; *count receives the total number of matches (may exceed cap).
hipError_t launch_near_grid(int ndof, const float* queries, int nq, long long qoff, const CellGrid& g,
                            const int* starts, const int* counts, const float* sorted, const int* sid, int n, float D2,
                            int* out_q, int* out_node, float* out_d2, int* count, int cap, hipStream_t st);
hipError_t launch_fill_u32(unsigned int* p, unsigned int v, long long n, hipStream_t st);

// Launch shape of the persistent separating-vector kernel: a kernel variant
// (block size, occupancy target in waves per SIMD), the pairs of one work
// unit, and the grid: every CU filled with as many blocks as are resident at
// once (the runtime's occupancy for the variant and the scene's LDS).
struct SvaShape {
  int block = 256;
  int waves = 5;
  int ppu = 32;
  int grid = 0;
  int gscene = 0;  // 1: the kernel reads the scene from the global image (beyond LDS)
};
// Variants: (256, 5), (256, 4), (1024, 4), each with the scene in LDS or,
// when the image and the block's rings do not fit, read from the global
// image (gscene, set here).  Fills in grid and gscene; an error for an
// unknown variant or ppu outside [1, 256].
hipError_t sva_shape_finish(const SceneImg& img, SvaShape* sh, int device);
// The same for the multi-scene kernel whose largest scene image has img_bytes.
hipError_t sva_multi_shape_finish(size_t img_bytes, SvaShape* sh, int device);
// The shape used until a caller (or the autotuner) sets another.
SvaShape sva_default_shape(const SceneImg& img);

}  // namespace gprm
