// planner.h -- device side of the single-process PRM (prm.cpp): the trees,
// their cell indices and the step kernels (planner.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "runtime.h"

namespace gprm {

// Counts shared by the planner's kernels and the host (device memory, one
// per solve).  Trees only grow; n, ecount and the flags are written by the
// insertion kernels, the totals by the step kernels.
struct PlanState {
  int n[2];                         // nodes per tree
  unsigned ecount[2];               // undirected edges stored per tree
  int conn;                         // 1 once a new node joined both trees (set once)
  int conn_left, conn_right;        // that node's index in tree 0 and tree 1
  int full;                         // a tree reached its node or edge capacity
  int over_bound;                   // an edge had more samples than the rounds planned (never expected)
  unsigned long long point_checks;  // candidates decided by the exact point check
  unsigned long long dropped;       // candidates dropped unchecked by the sampling bias
  unsigned long long edges_checked; // edges decided by the exact edge check
  unsigned long long edges_free;
  unsigned long long edge_samples;  // sum of testnum over checked edges
  unsigned long long edges_skipped; // candidate edges left unchecked by the edge budget
};

// Per-step counts of one worker (device, read back once per step).
struct StepInfo {
  int nnew;      // new nodes taken from the candidates
  int accepted;  // free, undropped candidates (before the per-tree quota)
  int E_all;     // candidate edges found by the radius search
  int E;         // min(E_all, nbuf): the edges of the step
  int S;         // edges in groups above the edge budget (scored by the model)
  int pad[3];
};

// One tree on the device.
struct PlanTree {
  float* q;        // SoA ndof x cap: component i of node m at q[m + cap i]
  int* surr;       // [cap] neighbours per node (the reference's surrnum)
  int* n;          // &PlanState::n[tree]
  double* cum;     // [cap] inclusive prefix sums of the base weights 1/(1+surr)^3
  double* bsum;    // [nb + 1] block sums of that scan (exclusive after the top pass), total at [nb]
  int* efrom;      // [ecap] undirected edges: new node,
  int* eto;        //        the node it joined,
  float* ew;       //        length
  unsigned* ecount;  // &PlanState::ecount[tree]
  int ecap;
  // cell index of the nodes [0, n): counting sort by cell
  int* cell_id;    // [cap]
  int* counts;     // [ncells]
  int* starts;     // [ncells]
  int* cursor;     // [ncells]
  int* cbsum;      // [cells blocks + 1] block sums of the count scan
  float* sorted;   // SoA ndof x cap: node coordinates in cell order
  int* sid;        // [cap] node of each sorted slot
};

// Items per block of the planner's scans (1024 threads x 8).
constexpr int kPlanScanItems = 8192;
inline int plan_scan_blocks(long long n) { return (int)((n + kPlanScanItems - 1) / kPlanScanItems); }

// Base weights of both trees: cum[m] = sum_{l <= m} 1/(1+surr[l])^3 over the
// nodes [0, n), and bsum[nb] = the total (nb = plan_scan_blocks(cap)).
hipError_t launch_plan_weights(PlanTree t0, PlanTree t1, int nb, hipStream_t st);
// Candidates: slot k (tree k & 1) draws its base node m with probability
// proportional to the weight (the reference's acceptance rule,
// prmsolver.cpp:2584-2624) and takes q_m plus a uniform offset in [-D, D)
// per coordinate; cand SoA with stride batch.
hipError_t launch_plan_sample(int ndof, int batch, PlanTree t0, PlanTree t1, int cap, int nb, float D,
                              unsigned long long key, float* cand, hipStream_t st);
// Sampling bias: skip[k] = (k & 2) && prob[k] > 0.5.
hipError_t launch_plan_bias_mask(const float* prob, int batch, unsigned char* skip, hipStream_t st);
// New nodes: in slot order, the free (code 0), unskipped candidates, at most
// want0 of tree 0 and want1 of tree 1, into qnew (SoA, stride num);
// info->nnew, info->accepted (info->E_all reset); state point_checks /
// dropped.  tile: 3 ints per 1024 candidates.
hipError_t launch_plan_select(int ndof, int batch, const int* codes, const unsigned char* skip, const float* cand,
                              int want0, int want1, int num, float* qnew, int* tile, StepInfo* info, PlanState* state,
                              hipStream_t st);
// Radius search of the step's new nodes in both trees, one group b = 2j + g
// per (new node j, tree g): every (node, squared distance, group) with
// distance^2 < D2 (the reference's get_near_vertices float sum) is a match;
// up to nbuf of them are stored (enode, ed2, egroup), cnt[b] = the group's
// stored edges, info->E_all = all matches, info->E = min(E_all, nbuf).
// With budget B > 0 a group keeps at most min(512, nbuf / groups) of its
// matches, its shortest (ties by traversal order), so a step never has
// more than nbuf; they are contiguous at off[b] (groups placed as their
// blocks finish), and sc_off = exclusive scan of the stored edges of the
// groups with more than B (the groups the model scores), info->S its
// total; sc_off has groups + 1 entries.  Without a budget the matches are
// in arrival order (off and sc_off unused).
hipError_t launch_plan_near(int ndof, int num, const float* qnew, StepInfo* info, PlanTree t0, PlanTree t1, int cap,
                            const CellGrid& g, float D2, int* cnt, int* off, int* sc_off, int nbuf, int budget,
                            int* enode, float* ed2, int* egroup, hipStream_t st);
// Shortest paths from node 0 in both trees (dist, parent: 2 x cap ints):
// init, then relaxation rounds over the edges (max_edges >= both trees'
// edge counts) that set changed[tree] when they lower a distance, then
// parents, the chains from the connection nodes back to the roots (chain:
// 2 x maxlen, len[2], -1 when broken) and their coordinates (coords AoS,
// 2 x maxlen x ndof).
hipError_t launch_plan_paths_init(PlanTree t0, PlanTree t1, int cap, int* dist, int* parent, hipStream_t st);
hipError_t launch_plan_paths_relax(PlanTree t0, PlanTree t1, int cap, int max_edges, int* dist, int* changed,
                                   hipStream_t st);
hipError_t launch_plan_paths_finish(int ndof, PlanTree t0, PlanTree t1, int cap, int max_edges, const int* dist,
                                    int* parent, const PlanState* state, int maxlen, int* chain, int* len, float* coords,
                                    hipStream_t st);
// Points scored along the edges of groups above the budget: edge e of such a
// group b, position p = sc_off[b] + e - off[b], points p*S8 + s at
// qa + t (qb - qa), t = (s+1)/(S8+1); out SoA with stride ns.
hipError_t launch_plan_score_points(int ndof, int E, int nbuf, int budget, int S8, const int* cnt, const int* off,
                                    const int* sc_off, const int* egroup, const int* enode, const float* qnew, int num,
                                    PlanTree t0, PlanTree t1, int cap, float* out, long long ns, hipStream_t st);
// Edge budget: in every group with more than B edges keep the B lowest by
// (score > 1/2, squared length, index), score = max of the edge's S8
// predictions (prob may be null: score 0): the shortest edges the model
// predicts free first, then the shortest predicted blocked; skip the
// others.  state edges_skipped.
hipError_t launch_plan_rank(int E, int nbuf, int budget, int S8, const int* cnt, const int* off, const int* sc_off,
                            const int* egroup, const float* ed2, const float* prob, float* score,
                            unsigned char* skip, PlanState* state, hipStream_t st);
// Endpoints of the step's edges, the lexicographically smaller
// configuration (first differing coordinate) as qs.
hipError_t launch_plan_gather(int ndof, int E, const int* egroup, const int* enode, const float* qnew, int num,
                              PlanTree t0, PlanTree t1, int cap, float* qs, float* qe, hipStream_t st);
// After the edge checks: free edges per group, state edges_checked /
// edges_free / edge_samples (total: the edge prologue's sample sum);
// over_bound if the prologue's largest testnum (max_t) exceeds maxT_bound.
hipError_t launch_plan_edges_done(int E, const int* egroup, const int* res, const unsigned char* skip, int* free_cnt,
                                  const unsigned long long* total, const int* max_t, int maxT_bound,
                                  PlanState* state, hipStream_t st);
// Insertion (calc_conn + part3a, prmsolver.cpp:2076-2187): every new node
// with a free edge to a tree joins it with all its free edges; the first
// node joining both trees ends the search (later nodes of the step are not
// inserted).  Then both cell indices are rebuilt.
hipError_t launch_plan_insert(int ndof, int num, int E, const StepInfo* info, const int* free_cnt, int* pos,
                              const float* qnew, const int* egroup, const int* enode, const float* ed2, const int* res,
                              PlanTree t0, PlanTree t1, int cap, PlanState* state, hipStream_t st);
hipError_t launch_plan_index(int ndof, PlanTree t0, PlanTree t1, int cap, const CellGrid& g, hipStream_t st);

}  // namespace gprm
