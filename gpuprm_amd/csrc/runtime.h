// runtime.h -- internal interface between the host runtime (gprm_api.cpp) and
// the HIP kernels (kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gprm {

constexpr int kFkBlock = 256;       // threads per block of the kinematics kernels
constexpr int kMaxIter = 10;        // collision4.h:17-18 (max_for_loop, max_for_loop_whole)
constexpr int kChunk = 64;          // items claimed per wave from the work counter
constexpr int kMaxDofDev = 16;

// Robot parameters (modified DH, robot.h:10-22) and joint bounds.
struct DevRobot {
  int ndof;
  float a[kMaxDofDev], alpha[kMaxDofDev], q[kMaxDofDev], d[kMaxDofDev];
  float mins[kMaxDofDev], maxs[kMaxDofDev];
  int types[kMaxDofDev];
};

// One contiguous device image of a scene, staged whole into LDS by the
// separating-vector kernel.  Sections (byte offsets, 16-byte aligned):
//   verts  float4[sumn]   vertex (x,y,z,0) in the polytope's own frame
//   pairs  int4[M]        (vertex base of P, vertex base of Q, frame of P, frame of Q)
//   vinfo  int[sumn]      (adjacency offset << 8) | adjacency count
//   adj    uint16[summ]   neighbour vertex ids (scene-global), rows as stored
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
  const float* frames;     // link frames, component-major: entry c of link s (1..ndof)
                           // of item i at frames[(12 (s-1) + c) fstride + i]
  long long fstride;       // item stride of frames (n of the kinematics launch, or its cap)
  float4* hist;            // [kMaxIter][grid*block] per-lane iteration history
  int nitems;              // item count, or -1: read *nitems_dev
  const int* nitems_dev;
  unsigned int* counter;   // work counter (zero before launch)
  int pairs_per_unit;      // pairs of one work unit (a lane's claim)
  int units_per_item;      // ceil(M / pairs_per_unit)
  // point mode
  int* word;               // [item] first colliding pair and code, see fk_points_kernel
  // edge mode
  const int* item_edge;    // [item] owning edge
  int* edge_res;           // [edge] nonzero once any sample collides
  // optional counters, added once per block at exit (may be null):
  // [0] pair tests ending separated, [1] pair tests ending in a collision,
  // [2] pair tests skipped by early exit, [3] units completed, [4] iterations
  // run, [5] pair tests decided separated by the bounding-sphere test
  unsigned long long* tally;
};
constexpr int kTallyCount = 6;

// Edge work description.
struct EdgeArgs {
  const float* qs; const float* qe; long long offset;  // SoA, component i of edge k at [k + offset*i]
  int nedges;
  const int* testnum;      // [edge]
  const int* stride;       // [edge] sample order: position j -> (j*stride) mod testnum
  int* edge_res;           // [edge]
};

size_t sva_lds_bytes(const SceneImg& img);
int sva_max_lds_bytes();
int sva_block_for(const SceneImg& img);

// Kinematics of n configurations; frames component-major with item stride n.
hipError_t launch_fk_points(const DevRobot& rb, const float* q, long long offset, int n, float* frames,
                            int* word, int check_bounds, hipStream_t st);
hipError_t launch_finalize_points(const int* word, int* codes, int n, hipStream_t st);
hipError_t launch_sva(const SvaArgs& a, int edge_mode, int grid, hipStream_t st);
hipError_t launch_edge_prologue(const DevRobot& rb, float dq, EdgeArgs ea, int* testnum_out,
                                int* stride_out, int* max_testnum, unsigned long long* total,
                                hipStream_t st);
hipError_t launch_edge_compact(EdgeArgs ea, int e0, int e1, int lo, int hi, int* item_edge,
                               int* item_pos, int* nitems, hipStream_t st);
// Kinematics of the *nitems (<= cap) compacted samples; frames component-major
// with item stride cap.
hipError_t launch_edge_fk(const DevRobot& rb, EdgeArgs ea, const int* item_edge, const int* item_pos,
                          const int* nitems, int cap, float* frames, hipStream_t st);
hipError_t launch_fill_i32(int* p, int v, long long n, hipStream_t st);
// Radius neighbours: every (query j, node l) with |node_l - query_j|^2 < D2,
// appended to (out_q, out_node, out_d2) up to cap entries; *count receives the
// total number of matches (may exceed cap).
hipError_t launch_near(int ndof, const float* queries, int nq, long long qoff, const float* nodes, int nn,
                       long long noff, float D2, int* out_q, int* out_node, float* out_d2, int* count, int cap,
                       hipStream_t st);
hipError_t launch_fill_u32(unsigned int* p, unsigned int v, long long n, hipStream_t st);

// Grid of the persistent separating-vector kernel: every CU filled as far as
// LDS and the wave limit allow.
int sva_grid(const SceneImg& img, int device);

}  // namespace gprm
