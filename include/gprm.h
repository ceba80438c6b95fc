/*
 * gprm.h -- C ABI of the MI355X collision-check path (libgprm.so).
 *
 * Drop-in for the configuration-space interface of GPUPRM: every entry point
 * below replaces one member of Configspace<ndof> / RobotConfigspace<ndof>
 * (paths relative to the reference repository root).  Layout conventions are
 * the reference's:
 *   - single configuration: AoS float[ndof];
 *   - batches and edges: SoA, component i of item k at q[k + offset*i]
 *     (prm/src/lib/prm/configspace.hpp:17-26);
 *   - results: int per item; 0 = free, nonzero = blocked.  Codes are the
 *     reference's: 1 out of bounds, l+1 no separating half plane found in
 *     iteration l, -1 iteration cap reached (collision4.h:296-331).
 * All entry points return GPRM_OK or a negative GPRM_E_* code; they never
 * abort.  A context is used by one host thread at a time (the reference's
 * contract, robotconfigspace.cpp:621-650); different contexts are independent.
 * Host arrays are owned by the caller and copied at call time; result arrays
 * of asynchronous requests are written by gprm_indicator2_wait.
 */
#ifndef GPRM_H
#define GPRM_H

#ifdef __cplusplus
extern "C" {
#endif

#define GPRM_OK 0
#define GPRM_E_INVALID (-1)   /* bad argument (null pointer, N < 0, bad index) */
#define GPRM_E_SCENE (-2)     /* scene could not be read or is malformed */
#define GPRM_E_HIP (-3)       /* HIP runtime error (no device, launch failure) */
#define GPRM_E_CAPACITY (-4)  /* scene or call beyond a stated limit (vertex count,
                                 degree, a multi-scene batch's LDS, the no-alloc
                                 maximum) */
#define GPRM_E_REQUEST (-5)   /* unknown or already completed request id */
#define GPRM_E_STATE (-6)     /* context not initialised (call gprm_init) */

typedef struct gprm_ctx gprm_ctx;

/* RobotConfigspace(std::string path, float dq, int nbuf)
 * (prm/src/lib/robotspace/robotconfigspace.h:42, .cpp:84-98; scene directory
 * format of load_config, robotconfigspace.cpp:100-208).  nbuf is the
 * reference's per-call edge capacity; here larger calls are split, so it is
 * only a sizing hint. */
int gprm_create_from_dir(const char* path, float dq, int nbuf, gprm_ctx** out);

/* RobotConfigspace(robot, polys, sys, N, from, to, M, mins, maxs, dq, nbuf)
 * (robotconfigspace.h:34-41), flattened: the N polytopes' vertices (3 floats
 * each), dsp/cnt (one per vertex, local offsets) and dest (local vertex ids)
 * back to back; n[k], m[k] their sizes.  Pairs (from, to) sorted by from
 * (polytope4data.h:225).  All arrays are copied. */
int gprm_create(int ndof, const float* a, const float* alpha, const float* q, const float* d,
                const int* types, int N, const int* sys, const int* n, const int* m,
                const float* vertices, const int* dsp, const int* cnt, const int* dest,
                const int* from, const int* to, int M, const float* mins, const float* maxs,
                float dq, int nbuf, gprm_ctx** out);

/* Configspace::init(ressource_rank, ressource_size) (configspace.hpp:16;
 * robotconfigspace.cpp:212-287): selects device rank*count/size, uploads the
 * scene. */
int gprm_init(gprm_ctx* ctx, int rank, int size);

/* Frees every device and host resource (the reference's clear() is
 * unimplemented, robotconfigspace.cpp:289-295). */
void gprm_destroy(gprm_ctx* ctx);

/* Configspace::indicator(const float* q) (configspace.hpp:22;
 * robotconfigspace.cpp:312-346): bounds check, then all pairs; *code gets the
 * reference's return value. */
int gprm_indicator1(gprm_ctx* ctx, const float* q, int* code);

/* Configspace::indicator(q, res, N, offset) (configspace.hpp:17-20; the
 * reference leaves it unimplemented, robotconfigspace.cpp:304-309): res[k]
 * gets indicator1 of item k. */
int gprm_indicator(gprm_ctx* ctx, const float* q, int* res, int N, int offset);

/* Configspace::indicator2(qs, qe, res, N, offset) (configspace.hpp:23-26;
 * robotconfigspace.cpp:616-752): res[k] != 0 iff some sample of edge k,
 * discretised as the reference does (testnum = (int)(|qe-qs|_2/dq)+1 samples,
 * robotconfigspace.cpp:660-673,409-420), is blocked.  No bounds check, as in
 * the reference. */
int gprm_indicator2(gprm_ctx* ctx, const float* qs, const float* qe, int* res, int N, int offset);

/* indicator2_async / indicator2_async_wait (configspace.hpp:27-28;
 * robotconfigspace.cpp:616-742).  Several requests may be in flight, each on
 * its own stream; res is written when wait returns. */
int gprm_indicator2_async(gprm_ctx* ctx, const float* qs, const float* qe, int* res, int N, int offset,
                          int* request);
int gprm_indicator2_wait(gprm_ctx* ctx, int request);

/* Compressed indicator2 (configspace.hpp:30-32; unimplemented in the
 * reference, robotconfigspace.cpp:766-772): for i < M and j < numqe[i],
 * res[posqe[i]+j] is the indicator2 result of the edge from qs item i to qe
 * item posqe[i]+j (qs has M items, stride offset; qe stride offset). */
int gprm_indicator2_compressed(gprm_ctx* ctx, const float* qs, int M, const float* qe, int* res,
                               const int* posqe, const int* numqe, int offset);

/* check_boundaries, batch and single (configspace.hpp:33-37;
 * robotconfigspace.cpp:779-802). */
int gprm_check_boundaries(gprm_ctx* ctx, const float* q, int* res, int N, int offset);
int gprm_check_boundaries1(gprm_ctx* ctx, const float* q, int* out);

/* min(i), max(i), deltaq(), dim() (configspace.hpp:39-47) and
 * get_numthreads_all() (robotconfigspace.h:112): samples checked by edge
 * queries since creation, counted as the reference counts them (sum of
 * testnum). */
float gprm_min(gprm_ctx* ctx, int i);
float gprm_max(gprm_ctx* ctx, int i);
float gprm_deltaq(gprm_ctx* ctx);
int gprm_dim(gprm_ctx* ctx);
long long gprm_numthreads_all(gprm_ctx* ctx);

/* Device-resident variants (no reference counterpart): the same computations
 * on arrays already in device memory, enqueued on `stream` (a hipStream_t,
 * NULL = the context's stream) and complete when the call returns.  For
 * callers that keep configurations in HBM (the planner's sample buffers). */
int gprm_indicator_dev(gprm_ctx* ctx, const float* d_q, int* d_res, int N, long long offset,
                       void* stream);
/* gprm_indicator_dev without any allocation (no reference counterpart; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same results as
 * gprm_indicator_dev.  Every buffer beyond d_q and d_res is a fixed scratch
 * array of 74.3 MiB that the library holds on each device it runs on until
 * the process ends.  N may be at most 65536 configurations for arms of up to
 * 8 joints (24 MiB of frames / (48 bytes x joints), rounded down to a
 * multiple of 64, beyond: 32768 at 16 joints); a larger N returns
 * GPRM_E_CAPACITY.  The call
 * only enqueues kernels on `stream` (NULL = the null stream) and returns;
 * d_res is complete when that stream reaches the end of the call.  Calls
 * share the device's scratch array: calls whose work overlaps on one GPU
 * race (serialise them, e.g. on one stream). */
int gprm_indicator_dev_noalloc(gprm_ctx* ctx, const float* d_q, int* d_res, int N, long long offset,
                               void* stream);
typedef struct {
  long long samples_total;     /* sum of testnum: the reference's work for the call */
  long long samples_scheduled; /* samples handed to the kernels after early exit (a
                                  round scheduled while the previous one runs
                                  may include samples of edges it blocks) */
  int max_testnum;
  int rounds;                  /* compaction rounds launched */
} gprm_edge_stats;
int gprm_indicator2_dev(gprm_ctx* ctx, const float* d_qs, const float* d_qe, int* d_res, int N,
                        long long offset, void* stream, gprm_edge_stats* stats);
/* Kinematics<ndof>::calculate(q, offset) (kinematics.h:28-37) for N
 * configurations in device memory (d_q[i + offset j], offset >= N): the
 * ndof + 1 frames it leaves in trafos, as the collision kernels use them,
 * to d_frames[(i (ndof + 1) + s) 12 + 3 c + r] (frame s, column c of four,
 * row r of x, y, z; trafo4, geo4.h:22-24). */
int gprm_kinematics_dev(gprm_ctx* ctx, const float* d_q, float* d_frames, int N, long long offset,
                        void* stream);

/* Kernel timing and work counters (no reference counterpart; the reference's
 * only counter is numthreads_all).  While enabled, every kinematics and
 * separating-vector launch is bracketed by HIP events on its own stream and
 * the kernels add their work counts; totals accumulate until read or
 * re-enabled.  Enabling resets the totals. */
typedef struct {
  long long sva_launches;      /* separating-vector kernel launches */
  double sva_ms;               /* their summed HIP-event durations */
  long long fk_launches;       /* kinematics kernel launches */
  double fk_ms;
  long long pairs_separated;   /* separating-vector tests that ended separated (code 0) */
  long long pairs_colliding;   /* ... that ended with a nonzero code */
  long long units_dropped;     /* pair tests skipped by early exit */
  long long units_done;        /* work units run to their end */
  long long iterations;        /* separating-vector iterations run */
  long long bound_separated;   /* pair tests whose first iteration is decided
                                  separated by a bounding-box margin test */
} gprm_profile_stats;
int gprm_profile_enable(gprm_ctx* ctx, int enable);
int gprm_profile_read(gprm_ctx* ctx, gprm_profile_stats* out);

/* Single-process bidirectional PRM on the device (the reference's
 * PRMSolver<ndof>, prm/src/lib/prm/prmsolver.{h,cpp}, process 5 restated for
 * one GPU: prmsolver.cpp:1765-2187).  Grows a tree from qstart and one from
 * qgoal; every node and edge entering a tree is decided by the exact checks
 * above.  Steps run on the device; `workers` steps are in flight at once
 * (process 5 keeps two: one worker's edge checks run while the other
 * worker's nodes are inserted and its next step is sampled), each read back
 * once.  Its edges are checked with the lexicographically smaller endpoint
 * first.  On success res->found = 1 and path (AoS, ndof floats per node, up
 * to path_cap nodes) runs from qstart to qgoal. */
typedef struct {
  int num;                  /* new nodes per step (the reference's --num), at most 16384 */
  float D;                  /* connection radius and sampling half-width (--distD) */
  int max_nodes;            /* capacity of each tree */
  int max_steps;            /* steps before giving up (--maxsteps) */
  unsigned long long seed;
  int nbuf;                 /* candidate edges per step and tree */
  void* surrogate;          /* optional model (gprm_surrogate_*), or NULL */
  int bias_sampling;        /* with a model: every second pair of candidates must be
                               predicted free (p < 0.5); the rest are dropped unchecked */
  int order_edges;          /* with a model and edge_budget: edges the model predicts
                               blocked (probability at the edge's midpoint above 1/2)
                               rank after the others */
  int edge_budget;          /* 0: check every candidate edge (the reference's rule);
                               k > 0: at most k per (new node, tree), shortest first
                               (among predicted-free edges first with order_edges),
                               chosen from the group's min(512, nbuf / (2 num))
                               shortest candidates */
  const char* output_dir;   /* NULL, or a directory that receives both trees and the
                               path in the reference's output layout (store_results,
                               prmsolver.cpp:2947-3055; prm/matlab/lib/prmoutputread.m) */
  int workers;              /* steps in flight: 1, or 2 as the reference's process 5
                               (prmsolver.cpp:1830-1861); 0 = 1 */
} gprm_prm_params;
typedef struct {
  int found, steps, path_len;
  long long nodes_left, nodes_right;
  long long point_checks;   /* candidate configurations checked exactly */
  long long edges_checked, edges_free;
  long long edge_samples;   /* sum of testnum over checked edges */
  long long model_predictions;
  long long candidates_skipped;  /* sampling candidates dropped unchecked by the bias */
  long long edges_skipped;       /* candidate edges left unchecked by edge_budget */
  double ms_total;               /* wall time of the call */
  /* device time of the steps' parts (HIP events, summed over steps; the
   * workers' parts overlap, so the sum may exceed ms_total): sampling with
   * the point checks, radius search, edge checks, insertion */
  double ms_sampling, ms_near, ms_edges, ms_insert;
  double ms_model;               /* model inference inside the steps (device time) */
  double ms_wait;                /* host time blocked on the steps' read-backs */
  double ms_final;               /* shortest paths on the device (Bellman-Ford rounds over
                                    the edge lists), the path read back; with output_dir
                                    also the trees */
} gprm_prm_result;
int gprm_prm_solve(gprm_ctx* ctx, const float* qstart, const float* qgoal, const gprm_prm_params* params,
                   gprm_prm_result* res, float* path, int path_cap);

/* Radius neighbours on the device (the planner's get_near_vertices,
 * prm/src/lib/prm/prmsolver.cpp:2532-2574): every pair (query j, node l) with
 * sum_i (node_l[i] - query_j[i])^2 < D^2, the sum taken in float over i in
 * order.  Nodes and queries are SoA device arrays (component i of item k at
 * [k + off*i], off >= count); mins/maxs (host, ndof) bound the nodes.  Up to
 * cap pairs go to the device arrays d_out_* in no particular order; *total
 * receives the number of pairs found (may exceed cap).  Through a cell index
 * over the first coordinates (up to five) instead of the reference's q[0]
 * blocks. */
int gprm_radius_neighbours(int ndof, const float* d_nodes, int n, long long noff, const float* d_queries, int nq,
                           long long qoff, const float* mins, const float* maxs, float D, int* d_out_q,
                           int* d_out_node, float* d_out_d2, int cap, long long* total, void* stream);

/* Learned collision-probability model (no reference counterpart): an MLP
 * 7 -> 256 -> 256 -> 256 -> 1 (SiLU, sigmoid output) on BF16 MFMA with FP32
 * accumulation, trained by its own forward/backward kernels and Adam on
 * labels of the exact checker (code != 0).  It only ranks configurations for
 * gprm_prm_solve; it never decides one.  Device arrays are SoA with the given
 * component offset; `batch` bounds n of a training call. */
int gprm_surrogate_create(int ndof, int batch, unsigned long long seed, void** out);
void gprm_surrogate_destroy(void* model);
int gprm_surrogate_param_count(void);
int gprm_surrogate_get_params(void* model, float* out);
int gprm_surrogate_set_params(void* model, const float* in);
int gprm_surrogate_gradients(void* model, const float* d_q, long long offset, const int* d_codes, int n,
                             float* grads_out, float* loss_out);
/* One Adam step on n <= batch rows; *loss_out gets the batch loss before the
 * step.  With loss_out NULL the step is only enqueued (on the null stream):
 * synchronise before reading the parameters. */
int gprm_surrogate_train_step(void* model, const float* d_q, long long offset, const int* d_codes, int n, float lr,
                              float* loss_out);
int gprm_surrogate_predict(void* model, const float* d_q, int n, long long offset, float* d_prob, void* stream);

/* Launch shape of the persistent separating-vector kernel (no reference
 * counterpart: the reference launches one thread per configuration,
 * robotconfigspace.cpp:718-723).  Variants (block, waves): (256, 5), (256, 4),
 * (1024, 4); a variant is available when its blocks fit the scene in LDS. */
typedef struct {
  int block;            /* threads per block */
  int waves;            /* occupancy target, waves per SIMD */
  int pairs_per_unit;   /* pairs of one work unit (a wave takes them for 64 items), 1..256 */
  int grid;             /* set by the library: resident blocks over all CUs */
} gprm_launch_shape;
int gprm_get_launch_shape(gprm_ctx* ctx, gprm_launch_shape* out);
int gprm_set_launch_shape(gprm_ctx* ctx, const gprm_launch_shape* in);
/* Autotuner: runs the given device workload (point checks of n
 * configurations when d_qe is NULL, else n edges; SoA with offset >= n) with
 * every available variant x pairs_per_unit in {32, 64, 128, 256}, `reps` timed runs
 * each after one warm-up, keeps the shape with the fastest median run and
 * returns it with that median wall time in ms. */
int gprm_autotune(gprm_ctx* ctx, const float* d_q, const float* d_qe, int n, long long offset, int reps,
                  gprm_launch_shape* best, float* ms_best);

/* Batches of scenes (no reference counterpart: the reference checks one scene
 * per RobotConfigspace object and launch).  A batch runs the point checks of
 * many initialised contexts (same device, same ndof; they must outlive the
 * batch) as one sequence: one kinematics launch over all scenes, one
 * persistent separating-vector launch per vertex-count bucket, one finalize
 * launch.  Buckets: scenes with [0,512), [512,1024), [1024,2048) and >= 2048
 * vertices in total, each with its own launch shape. */
typedef struct gprm_batch gprm_batch;
typedef struct {
  int min_vertices, max_vertices;  /* the bucket's range of scene vertex counts */
  int scenes;                      /* scenes of the batch in it */
  int lds_bytes;                   /* largest scene image among them */
  gprm_launch_shape shape;
} gprm_bucket_info;
int gprm_batch_create(gprm_ctx* const* ctxs, int nscenes, gprm_batch** out);
void gprm_batch_destroy(gprm_batch* batch);
int gprm_batch_bucket_info(gprm_batch* batch, int bucket, gprm_bucket_info* out);   /* bucket 0..3 */
int gprm_batch_set_bucket_shape(gprm_batch* batch, int bucket, const gprm_launch_shape* shape);
/* Scene s: n[s] (<= 2^22) configurations at d_q[s] (device, SoA with
 * offset[s] >= n[s]) -> codes in d_res[s] (device), as gprm_indicator_dev
 * computes them.  use_graph: the sequence is captured into a hipGraph the
 * first time these arrays and sizes are seen and replayed while they stay the
 * same.  done_ms (host, nscenes entries, may be NULL): when each scene's last
 * check finished, in ms after the first separating-vector block started
 * (device clock).  Returns when the results are complete. */
int gprm_batch_indicator_dev(gprm_batch* batch, const float* const* d_q, const long long* offset, const int* n,
                             int* const* d_res, int use_graph, float* done_ms);
/* Edge checks of many scenes: scene s has n[s] (<= 2^22) edges, start and
 * end configurations at d_qs[s], d_qe[s] (device, SoA with offset[s] >= n[s])
 * -> d_res[s] (device), as gprm_indicator2_dev computes them (nonzero when a
 * sample collides; edges shorter than dq are blocked, robotconfigspace.cpp:
 * 411-413).  One prologue launch over all scenes' edges, read back once, then
 * launch steps over order-position windows: one compaction and one
 * kinematics launch over every scene per step, one separating-vector launch
 * per bucket.  use_graph: the steps are captured into a hipGraph and
 * replayed while the arrays, sizes and the scenes' largest testnums stay the
 * same.  done_ms as for gprm_batch_indicator_dev; samples_total (may be
 * NULL): the sum of testnum over all edges, the reference's sample count. */
int gprm_batch_indicator2_dev(gprm_batch* batch, const float* const* d_qs, const float* const* d_qe,
                              const long long* offset, const int* n, int* const* d_res, int use_graph, float* done_ms,
                              long long* samples_total);
/* Autotuner: for every bucket in use, times its separating-vector launch
 * (HIP events; reps runs after one warm-up, their median kept) on the given
 * workload with every available variant x pairs_per_unit in {32, 64, 128, 256}
 * and keeps the shape with the fastest median. */
int gprm_batch_autotune(gprm_batch* batch, const float* const* d_q, const long long* offset, const int* n,
                        int* const* d_res, int reps);
const char* gprm_batch_last_error(gprm_batch* batch);  /* NULL: error of the last create */

/* Last error message of the context (or of creation when ctx is NULL). */
const char* gprm_last_error(gprm_ctx* ctx);

/* Scene generators of the benchmark workloads (1: 4-DOF boxes, 2: 6-DOF
 * hulls, 3: 7-DOF hulls, 5: 7-DOF with random obstacle/hull counts, 6: 7-DOF
 * among 256 obstacles of 32 vertices, a scene image beyond LDS), written in
 * the reference's directory format. */
int gprm_scene_generate(int workload, unsigned long long seed, const char* out_dir);

#ifdef __cplusplus
}
#endif
#endif /* GPRM_H */
