/*
 * oracle.h -- CPU restatement of GPUPRM's exact collision-check path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg).  It is never linked into, loaded by or called from the product
 * (gpuprm_amd/), which must fail loudly when its HIP extension is missing.
 *
 * Every function restates one reference function, with the reference file:line
 * it follows (paths relative to the reference repository root):
 *   geometry        prm/src/lib/collision4/geo4.h
 *   separating vec. prm/src/lib/collision4/collision4.h
 *   kinematics      prm/src/lib/robotspace/kinematics.h
 *   scene packing   prm/src/lib/collision4/polytope4data.h:296-366
 *   scene format    prm/src/lib/robotspace/robotconfigspace.cpp:100-208
 *   indicator       prm/src/lib/robotspace/robotconfigspace.cpp:312-346
 *   edge host loop  prm/src/lib/robotspace/robotconfigspace.cpp:378-465,616-705
 *   bounds          prm/src/lib/robotspace/robotconfigspace.cpp:779-802
 *
 * Floating point: compiled with -ffp-contract=off -fno-fast-math (no FMA
 * contraction, IEEE division/sqrt), exactly like the reference's host build
 * (prm/makefile:93-94: g++ -O3, no -march).  Trigonometry calls glibc
 * sinf/cosf, which is what the reference's unqualified sin/cos on float
 * arguments bind to (collision4/cuda_head.h:25-28, geo4.h:184-187).
 *
 * Pinning (what makes this restatement trustworthy; see tests/):
 *   - golden vectors written by the reference's own CPU path compiled from its
 *     sources (oracle/Makefile -> oracle/_ref), committed under tests/golden/:
 *     FK floats, per-pair codes and iteration counts, indicator codes, edge
 *     flags, on the shipped config1 scene and generated scenes;
 *   - the reference's known answers indicator((0,0,0,0)) == 0 and
 *     indicator((3,3,3,3)) == 0 on config1 (prm/src/main.cpp:128-138);
 *   - closed forms (FK at q = 0, +-pi/2), brute-force separation of boxes and
 *     hulls by double-precision LP on tiny inputs.
 */
#ifndef GPRM_ORACLE_H
#define GPRM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAXDOF 16

typedef struct { float x, y, z, w; } ofloat4;
typedef struct { float x, y; } ofloat2;
/* 3x4 affine transform stored as 4 columns (geo4.h:22-24). */
typedef struct { ofloat4 col[4]; } otrafo4;

typedef struct {
  int ndof;
  float a[ORACLE_MAXDOF], alpha[ORACLE_MAXDOF], q[ORACLE_MAXDOF], d[ORACLE_MAXDOF];
  int types[ORACLE_MAXDOF];            /* 0 = rotational, 1 = prismatic (robot.h:6-7) */
  float mins[ORACLE_MAXDOF], maxs[ORACLE_MAXDOF];
  float dq;
  /* polytopes, packed as polytope4data (polytope4data.h:19-64) */
  int N, sumn, summ;
  int *sys, *n, *m, *dspn, *dspm;      /* [N] */
  ofloat4 *vertices;                   /* [sumn], w = 1 */
  int *dsp, *cnt;                      /* [sumn], local to the polytope */
  int *dest;                           /* [summ], local vertex ids */
  /* collision pairs as CRS over "from" (polytope4data.h:343-357) */
  int M;
  int *pair_dsp, *pair_cnt;            /* [N] */
  int *pair_dest;                      /* [M] */
} oracle_scene;

/* Build a scene from flat arrays.  vertices holds 3*sum(n) floats, dsp/cnt
 * sum(n) ints, dest sum(m) ints, all polytopes back to back.  Returns NULL on
 * invalid input.  from[] must be sorted ascending (polytope4data.h:225). */
oracle_scene* oracle_scene_create(int ndof, const float* a, const float* alpha,
                                  const float* q, const float* d, const int* types,
                                  const float* mins, const float* maxs, float dq,
                                  int N, const int* sys, const int* n, const int* m,
                                  const float* vertices, const int* dsp, const int* cnt,
                                  const int* dest, int M, const int* from, const int* to);
/* Load the reference's on-disk scene directory (robotconfigspace.cpp:100-208). */
oracle_scene* oracle_scene_load_dir(const char* path, float dq);
void oracle_scene_free(oracle_scene* s);

/* Forward kinematics (kinematics.h:27-37): T[0..ndof]. */
void oracle_fk(const oracle_scene* s, const float* q, otrafo4* T);

/* Separating-vector algorithm between polytopes k0 (P) and k1 (Q)
 * (collision4.h:227-332).  Returns 0 (separated), l+1 (no half plane) or -1
 * (iteration cap); *iterations receives l+1 as the reference sets it. */
int oracle_sva(const oracle_scene* s, int k0, int k1, const otrafo4* tp, const otrafo4* tq,
               int* iterations);

/* Pair loop without bounds check, as run per sample by kernel_indicator2
 * (robotconfigspace.cpp:423-453).  Optionally reports how many pairs ran. */
int oracle_check_config(const oracle_scene* s, const float* q, int* pairs_run);

/* RobotConfigspace::indicator(q) (robotconfigspace.cpp:312-346). */
int oracle_indicator1(const oracle_scene* s, const float* q);
/* Batched indicator with the contract of configspace.hpp:17-20 (SoA, item k
 * component i at q[k + offset*i]); each item gets indicator1 semantics (the
 * reference leaves this unimplemented, robotconfigspace.cpp:304-309). */
int oracle_indicator_batch(const oracle_scene* s, const float* q, int* res, int N, int offset,
                           int nthreads);

/* check_boundaries, batch and single (robotconfigspace.cpp:779-802). */
int oracle_check_boundaries(const oracle_scene* s, const float* q, int* res, int N, int offset);
int oracle_check_boundaries1(const oracle_scene* s, const float* q);

/* Edge prologue (robotconfigspace.cpp:660-673): testnum, testpos; returns
 * numthreads = sum(testnum). */
long long oracle_edge_testnum(const oracle_scene* s, const float* qs, const float* qe, int N,
                              int offset, int* testnum, long long* testpos);
/* indicator2 host loop (robotconfigspace.cpp:389-465,694-705): every sample of
 * every edge is checked; res[k] = code of the last colliding sample of edge k
 * (0 if none).  Returns numthreads. */
long long oracle_indicator2(const oracle_scene* s, const float* qs, const float* qe, int* res,
                            int N, int offset, int nthreads);
/* q of sample i of edge k (robotconfigspace.cpp:409-420). */
void oracle_edge_sample(const oracle_scene* s, const float* qs, const float* qe, int k,
                        int offset, int j, int testnum, float* q);

/* The pair loop of one configuration, recording for every pair test run (in
 * CRS order, stopping at the first nonzero code like the reference) its code
 * and, in the counting build, its FP32 operation count.  Returns the number
 * of pair tests run. */
int oracle_pair_profile(const oracle_scene* s, const float* q, int* codes, unsigned long long* flops);

/* Algorithmic FP32 operation counter (add/sub/mul/div/sqrt; negations and
 * compares excluded; sin/cos counted separately).  Non-zero only in the
 * ORACLE_COUNT build (liboracle_count.so). */
void oracle_counters(unsigned long long* flops, unsigned long long* trig);
void oracle_counters_reset(void);

#ifdef __cplusplus
}
#endif
#endif
