/*
 * oracle.c -- CPU restatement of GPUPRM's exact collision-check path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h for scope, citations and pinning).
 *
 * Every arithmetic operation is written out in the reference's evaluation
 * order through the counted helpers below, so that the same source gives both
 * the checker (liboracle.so) and the algorithmic-FLOP counter
 * (liboracle_count.so, -DORACLE_COUNT).  Build: -O2 -ffp-contract=off
 * -fno-fast-math.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* collision4.h:17-20 */
#define MAX_FOR_LOOP 10
#define MAX_FOR_LOOP_WHOLE 10
#define MAX_VERTICES_NUMBER 100
#define COLLISION_EPS 0.000001

#ifdef ORACLE_COUNT
static _Thread_local unsigned long long tl_flops, tl_trig;
static unsigned long long g_flops, g_trig;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
#define CNT() (++tl_flops)
#define CNT_TRIG() (++tl_trig)
static void flush_counters(void) {
  pthread_mutex_lock(&g_mu);
  g_flops += tl_flops; g_trig += tl_trig;
  tl_flops = tl_trig = 0;
  pthread_mutex_unlock(&g_mu);
}
void oracle_counters(unsigned long long* f, unsigned long long* t) {
  flush_counters();
  *f = g_flops; *t = g_trig;
}
void oracle_counters_reset(void) {
  pthread_mutex_lock(&g_mu);
  g_flops = g_trig = 0; tl_flops = tl_trig = 0;
  pthread_mutex_unlock(&g_mu);
}
#else
#define CNT() ((void)0)
#define CNT_TRIG() ((void)0)
static void flush_counters(void) {}
void oracle_counters(unsigned long long* f, unsigned long long* t) { *f = 0; *t = 0; }
void oracle_counters_reset(void) {}
#endif

static inline float A(float a, float b) { CNT(); return a + b; }
static inline float S(float a, float b) { CNT(); return a - b; }
static inline float Mu(float a, float b) { CNT(); return a * b; }
static inline float D(float a, float b) { CNT(); return a / b; }
static inline float Sq(float a) { CNT(); return sqrtf(a); }
/* `1.0/x` with a float x: double division rounded to float equals the
 * correctly rounded float quotient (53 >= 2*24+2), geo4.h:81,167. */
static inline float Rcp(float a) { CNT(); return (float)(1.0 / (double)a); }
static inline float fsin(float x) { CNT_TRIG(); return sinf(x); }
static inline float fcos(float x) { CNT_TRIG(); return cosf(x); }

/* ---------------- geo4 (geo4.h) ---------------- */

/* geo4.h:66-68 */
static inline float cross2(ofloat2 v, ofloat2 w) { return S(Mu(v.x, w.y), Mu(v.y, w.x)); }
/* geo4.h:80-83 (no zero guard: may produce inf/NaN on purpose) */
static inline void normalize2(ofloat2* u) {
  float factor = Rcp(Sq(A(Mu(u->x, u->x), Mu(u->y, u->y))));
  u->x = Mu(u->x, factor); u->y = Mu(u->y, factor);
}
/* geo4.h:130-132 */
static inline float sprod(const ofloat4* u, const ofloat4* v) {
  return A(A(Mu(u->x, v->x), Mu(u->y, v->y)), Mu(u->z, v->z));
}
/* geo4.h:165-171 */
static inline float normalize4(ofloat4* u) {
  float norm = Sq(A(A(Mu(u->x, u->x), Mu(u->y, u->y)), Mu(u->z, u->z)));
  float factor = Rcp(norm);
  if (factor == INFINITY) factor = 0.0f;
  u->x = Mu(u->x, factor); u->y = Mu(u->y, factor); u->z = Mu(u->z, factor);
  return norm;
}

/* trafo4::set (geo4.h:194-203) */
static void trafo_set(otrafo4* t, float a, float alpha, float q, float d) {
  float cq = fcos(q), sq = fsin(q), ca = fcos(alpha), sa = fsin(alpha);
  t->col[0].x = cq;          t->col[1].x = -sq;         t->col[2].x = 0.0f; t->col[3].x = a;
  t->col[0].y = Mu(sq, ca);  t->col[1].y = Mu(cq, ca);  t->col[2].y = -sa;  t->col[3].y = Mu(-sa, d);
  t->col[0].z = Mu(sq, sa);  t->col[1].z = Mu(cq, sa);  t->col[2].z = ca;   t->col[3].z = Mu(ca, d);
}
/* trafo4::apply(u, Tu) (geo4.h:215-221) */
static inline void t_apply(const otrafo4* t, const ofloat4* u, ofloat4* r) {
  const ofloat4* c = t->col;
  r->x = A(A(A(Mu(c[0].x, u->x), Mu(c[1].x, u->y)), Mu(c[2].x, u->z)), c[3].x);
  r->y = A(A(A(Mu(c[0].y, u->x), Mu(c[1].y, u->y)), Mu(c[2].y, u->z)), c[3].y);
  r->z = A(A(A(Mu(c[0].z, u->x), Mu(c[1].z, u->y)), Mu(c[2].z, u->z)), c[3].z);
}
/* trafo4::apply(u) in place (geo4.h:206-213) */
static inline void t_apply_ip(const otrafo4* t, ofloat4* u) {
  ofloat4 v = *u; t_apply(t, &v, u);
}
/* trafo4::apply_rot(u) in place (geo4.h:223-230) */
static inline void t_apply_rot_ip(const otrafo4* t, ofloat4* u) {
  const ofloat4* c = t->col;
  float a = u->x, b = u->y, cc = u->z;
  u->x = A(A(Mu(c[0].x, a), Mu(c[1].x, b)), Mu(c[2].x, cc));
  u->y = A(A(Mu(c[0].y, a), Mu(c[1].y, b)), Mu(c[2].y, cc));
  u->z = A(A(Mu(c[0].z, a), Mu(c[1].z, b)), Mu(c[2].z, cc));
}
/* trafo4::apply_rot_inv(u, Rtu) (geo4.h:249-255) */
static inline void t_apply_rot_inv(const otrafo4* t, const ofloat4* u, ofloat4* r) {
  const ofloat4* c = t->col;
  r->x = A(A(Mu(c[0].x, u->x), Mu(c[0].y, u->y)), Mu(c[0].z, u->z));
  r->y = A(A(Mu(c[1].x, u->x), Mu(c[1].y, u->y)), Mu(c[1].z, u->z));
  r->z = A(A(Mu(c[2].x, u->x), Mu(c[2].y, u->y)), Mu(c[2].z, u->z));
}
/* trafo4::lapply(T): this = T*this (geo4.h:272-279) */
static inline void t_lapply(otrafo4* self, const otrafo4* T) {
  t_apply_rot_ip(T, &self->col[0]);
  t_apply_rot_ip(T, &self->col[1]);
  t_apply_rot_ip(T, &self->col[2]);
  t_apply_ip(T, &self->col[3]);
}

/* ---------------- kinematics (kinematics.h:27-37) ---------------- */
void oracle_fk(const oracle_scene* s, const float* q, otrafo4* T) {
  memset(T, 0, sizeof(otrafo4) * (size_t)(s->ndof + 1));
  trafo_set(&T[0], 0.0f, 0.0f, 0.0f, 0.0f);
  for (int i = 0; i < s->ndof; ++i) {
    float qset = s->q[i], dset = s->d[i];
    if (s->types[i] == 0) qset = q[i];
    else dset = q[i]; /* q[i*offset], offset is always 1 */
    trafo_set(&T[i + 1], s->a[i], s->alpha[i], qset, dset);
    t_lapply(&T[i + 1], &T[i]);
  }
}

/* ---------------- collision4 (collision4.h) ---------------- */

typedef struct {
  const ofloat4* v; const int* dsp; const int* cnt; const int* dest; int n;
} opoly;

static inline opoly get_poly(const oracle_scene* s, int k) {
  opoly p;
  p.v = s->vertices + s->dspn[k];
  p.dsp = s->dsp + s->dspn[k];
  p.cnt = s->cnt + s->dspn[k];
  p.dest = s->dest + s->dspm[k];
  p.n = s->n[k];
  return p;
}

/* support_vertex_searcher::search_support_vertex (collision4.h:66-103):
 * first-improvement ascent over the vertex adjacency, in stored order. */
static inline void search_support(const opoly* P, int* p, const ofloat4* Sp) {
  float mx = sprod(&P->v[*p], Sp);
  for (;;) {
    int cont = 0;
    int imax = P->dsp[*p] + P->cnt[*p];
    for (int i = P->dsp[*p]; i < imax; ++i) {
      int v = P->dest[i];
      float dp = sprod(&P->v[v], Sp);
      if (dp > mx) { mx = dp; *p = v; cont = 1; break; }
    }
    if (!cont) break;
  }
}

/* orthtrafo23 (collision4.h:111-134) */
typedef struct { float M[6]; } orth23;
static inline orth23 orth_make(const ofloat4* rk) {
  orth23 o;
  float x = rk->x, y = rk->y, z = rk->z;
  float d2 = A(Mu(x, x), Mu(y, y));
  float d = Sq(d2);
  o.M[0] = D(Mu(x, z), d); o.M[1] = D(Mu(y, z), d); o.M[2] = -d;
  o.M[3] = D(-y, d);       o.M[4] = D(x, d);        o.M[5] = 0.0f;
  return o;
}
static inline ofloat2 mult23(const orth23* o, const ofloat4* x) {
  ofloat2 r;
  r.x = A(A(Mu(o->M[0], x->x), Mu(o->M[1], x->y)), Mu(o->M[2], x->z));
  r.y = A(A(Mu(o->M[3], x->x), Mu(o->M[4], x->y)), Mu(o->M[5], x->z));
  return r;
}
static inline void mult23T(const orth23* o, ofloat2 x, ofloat4* r) {
  r->x = A(Mu(o->M[0], x.x), Mu(o->M[3], x.y));
  r->y = A(Mu(o->M[1], x.x), Mu(o->M[4], x.y));
  r->z = A(Mu(o->M[2], x.x), Mu(o->M[5], x.y));
}

/* find_half_plane (collision4.h:136-223) */
static int find_half_plane(const ofloat4* R, int k, ofloat4* w, const ofloat4* rk) {
  float dp = sprod(w, rk);
  if (dp > 0) return 1;
  orth23 M = orth_make(rk);
  ofloat2 ra = mult23(&M, &R[0]);
  ofloat2 rb = mult23(&M, &R[1]);
  ofloat2 T;
  if (cross2(ra, rb) < 0) { ofloat2 t = ra; ra = rb; rb = t; }
  for (int i = 2; i < k; ++i) {
    T = mult23(&M, &R[i]);
    float rbcross = cross2(rb, T);
    if (cross2(ra, T) > 0) {
      if (rbcross > 0) rb = T;
    } else {
      if (rbcross < 0) ra = T;
      else return 0;
    }
  }
  normalize2(&ra);
  normalize2(&rb);
  T.x = A(ra.x, rb.x); T.y = A(ra.y, rb.y);
  normalize2(&T);
  mult23T(&M, T, w);
  return 1;
}

/* seperating_vector_algorithm (collision4.h:227-332) */
static int sva(const opoly* P, const opoly* Q, const otrafo4* tp, const otrafo4* tq,
               int* iterations) {
  int p = 0, q = 0;
  ofloat4 Sv, Pq, Pp;
  t_apply(tq, &Q->v[0], &Pq);
  t_apply(tp, &P->v[0], &Pp);
  Sv.x = S(Pq.x, Pp.x); Sv.y = S(Pq.y, Pp.y); Sv.z = S(Pq.z, Pp.z); Sv.w = 0.0f;
  float norm = normalize4(&Sv);
  if ((double)norm < COLLISION_EPS) Sv.x = 1.0f;

  /* w is read uninitialised by the reference only on the all-NaN path
   * (testnum == 1 edges), where the outcome does not depend on it. */
  ofloat4 w = {0.0f, 0.0f, 0.0f, 0.0f};
  ofloat4 R[MAX_FOR_LOOP];
  int combsp[MAX_FOR_LOOP], combsq[MAX_FOR_LOOP];
  int k, l;
  for (k = 0, l = 0; k < MAX_FOR_LOOP && l < MAX_FOR_LOOP_WHOLE; ++l) {
    ofloat4* rk = &R[k];
    ofloat4 Sp, Sqv, p0, q0, Sn;
    t_apply_rot_inv(tp, &Sv, &Sp);
    Sn.x = -Sv.x; Sn.y = -Sv.y; Sn.z = -Sv.z; Sn.w = Sv.w;   /* S*=-1.0 */
    t_apply_rot_inv(tq, &Sn, &Sqv);
    search_support(P, &p, &Sp);
    search_support(Q, &q, &Sqv);
    t_apply(tp, &P->v[p], &p0);
    t_apply(tq, &Q->v[q], &q0);
    rk->x = S(q0.x, p0.x); rk->y = S(q0.y, p0.y); rk->z = S(q0.z, p0.z); rk->w = 0.0f;
    normalize4(rk);
    float dp = sprod(&Sv, rk);
    if ((double)dp >= -COLLISION_EPS) {
      if (iterations) *iterations = l + 1;
      return 0;
    }
    combsp[k] = p;
    combsq[k] = q;
    int repetition = 0;
    for (int j = 0; j < k; ++j)
      if (combsp[j] == combsp[k] && combsq[j] == combsq[k]) { repetition = 1; break; }
    if (repetition) {
      Sv = w;
    } else {
      if (k == 1) {
        w.x = A(R[0].x, rk->x); w.y = A(R[0].y, rk->y); w.z = A(R[0].z, rk->z);
        normalize4(&w);
      }
      if (k >= 2 && !find_half_plane(R, k, &w, rk)) {
        if (iterations) *iterations = l + 1;
        return l + 1;
      }
      /* lin(S, -2.0*dp, rk, S): b = (float)(-2.0*dp) is exact */
      float b = -2.0f * dp;
      Sv.x = A(Sv.x, Mu(b, rk->x));
      Sv.y = A(Sv.y, Mu(b, rk->y));
      Sv.z = A(Sv.z, Mu(b, rk->z));
      ++k;
    }
  }
  if (iterations) *iterations = l + 1;
  return -1;
}

int oracle_sva(const oracle_scene* s, int k0, int k1, const otrafo4* tp, const otrafo4* tq,
               int* iterations) {
  opoly P = get_poly(s, k0), Q = get_poly(s, k1);
  int r = sva(&P, &Q, tp, tq, iterations);
  flush_counters();
  return r;
}

/* pair loop (robotconfigspace.cpp:320-345 and 430-453) */
static int check_config(const oracle_scene* s, const float* q, int* pairs_run) {
  otrafo4 T[ORACLE_MAXDOF + 1];
  oracle_fk(s, q, T);
  int run = 0;
  for (int k0 = 0; k0 < s->N; ++k0) {
    opoly P = get_poly(s, k0);
    const int* dest = s->pair_dest + s->pair_dsp[k0];
    int destnum = s->pair_cnt[k0];
    for (int l = 0; l < destnum; ++l) {
      int k1 = dest[l];
      opoly Q = get_poly(s, k1);
      ++run;
      int result = sva(&P, &Q, &T[s->sys[k0]], &T[s->sys[k1]], NULL);
      if (result != 0) {
        if (pairs_run) *pairs_run = run;
        return result;
      }
    }
  }
  if (pairs_run) *pairs_run = run;
  return 0;
}

int oracle_check_config(const oracle_scene* s, const float* q, int* pairs_run) {
  int r = check_config(s, q, pairs_run);
  flush_counters();
  return r;
}

int oracle_pair_profile(const oracle_scene* s, const float* q, int* codes, unsigned long long* flops) {
  otrafo4 T[ORACLE_MAXDOF + 1];
  oracle_fk(s, q, T);
  int run = 0;
  for (int k0 = 0; k0 < s->N; ++k0) {
    opoly P = get_poly(s, k0);
    const int* dest = s->pair_dest + s->pair_dsp[k0];
    for (int l = 0; l < s->pair_cnt[k0]; ++l) {
      opoly Q = get_poly(s, dest[l]);
#ifdef ORACLE_COUNT
      unsigned long long f0 = tl_flops;
#endif
      int result = sva(&P, &Q, &T[s->sys[k0]], &T[s->sys[dest[l]]], NULL);
#ifdef ORACLE_COUNT
      flops[run] = tl_flops - f0;
#else
      flops[run] = 0;
#endif
      codes[run++] = result;
      if (result != 0) {
        flush_counters();
        return run;
      }
    }
  }
  flush_counters();
  return run;
}

/* robotconfigspace.cpp:795-802 */
int oracle_check_boundaries1(const oracle_scene* s, const float* q) {
  for (int i = 0; i < s->ndof; ++i) {
    float qi = q[i];
    if (qi < s->mins[i] || qi > s->maxs[i]) return 1;
  }
  return 0;
}

/* robotconfigspace.cpp:779-791 */
int oracle_check_boundaries(const oracle_scene* s, const float* q, int* res, int N, int offset) {
  for (int ix = 0; ix < N; ++ix) {
    int r = 0;
    for (int i = 0; i < s->ndof; ++i) {
      float qi = q[ix + i * offset];
      if (qi < s->mins[i] || qi > s->maxs[i]) { r = 1; break; }
    }
    res[ix] = r;
  }
  return 0;
}

/* robotconfigspace.cpp:312-346 */
int oracle_indicator1(const oracle_scene* s, const float* q) {
  if (oracle_check_boundaries1(s, q) == 1) return 1;
  return oracle_check_config(s, q, NULL);
}

typedef struct {
  const oracle_scene* s; const float* q; int* res; int N, offset, t, nt;
} batch_job;

static void* batch_worker(void* arg) {
  batch_job* j = (batch_job*)arg;
  long long lo = (long long)j->N * j->t / j->nt, hi = (long long)j->N * (j->t + 1) / j->nt;
  float qq[ORACLE_MAXDOF];
  for (long long k = lo; k < hi; ++k) {
    for (int i = 0; i < j->s->ndof; ++i) qq[i] = j->q[k + (long long)i * j->offset];
    if (oracle_check_boundaries1(j->s, qq) == 1) { j->res[k] = 1; continue; }
    j->res[k] = check_config(j->s, qq, NULL);
  }
  flush_counters();
  return NULL;
}

int oracle_indicator_batch(const oracle_scene* s, const float* q, int* res, int N, int offset,
                           int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  batch_job jobs[256];
  for (int t = 0; t < nthreads; ++t) {
    batch_job b = {s, q, res, N, offset, t, nthreads};
    jobs[t] = b;
    if (nthreads == 1) batch_worker(&jobs[0]);
    else pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  return 0;
}

/* robotconfigspace.cpp:660-673 */
long long oracle_edge_testnum(const oracle_scene* s, const float* qs, const float* qe, int N,
                              int offset, int* testnum, long long* testpos) {
  long long numthreads = 0;
  for (int l = 0; l < N; ++l) {
    if (testpos) testpos[l] = numthreads;
    float norm = 0.0f;
    for (int i = 0; i < s->ndof; ++i) {
      float diff = qe[l + offset * i] - qs[l + offset * i];
      norm += diff * diff;
    }
    norm = sqrtf(norm);
    testnum[l] = (int)(norm / s->dq) + 1;
    numthreads += testnum[l];
  }
  return numthreads;
}

/* robotconfigspace.cpp:409-420 */
void oracle_edge_sample(const oracle_scene* s, const float* qs, const float* qe, int k,
                        int offset, int j, int testnum, float* q) {
  float c1, c2;
  if (testnum > 0) {
    c1 = (float)j / (float)(testnum - 1);
    c2 = (float)(1.0 - (double)c1);
  } else {
    c1 = 1.0f;
    c2 = 0.0f;
  }
  for (int i = 0; i < s->ndof; ++i)
    q[i] = A(Mu(c1, qs[k + offset * i]), Mu(c2, qe[k + offset * i]));
}

typedef struct {
  const oracle_scene* s; const float* qs; const float* qe; int* res; const int* testnum;
  int N, offset, t, nt;
} edge_job;

static void* edge_worker(void* arg) {
  edge_job* j = (edge_job*)arg;
  long long lo = (long long)j->N * j->t / j->nt, hi = (long long)j->N * (j->t + 1) / j->nt;
  float q[ORACLE_MAXDOF];
  for (long long k = lo; k < hi; ++k) {
    int r = 0;
    /* host loop visits samples in increasing thread index; the last nonzero
     * code written wins (robotconfigspace.cpp:398,458) */
    for (int i = 0; i < j->testnum[k]; ++i) {
      oracle_edge_sample(j->s, j->qs, j->qe, (int)k, j->offset, i, j->testnum[k], q);
      int code = check_config(j->s, q, NULL);
      if (code != 0) r = code;
    }
    j->res[k] = r;
  }
  flush_counters();
  return NULL;
}

long long oracle_indicator2(const oracle_scene* s, const float* qs, const float* qe, int* res,
                            int N, int offset, int nthreads) {
  int* testnum = (int*)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
  long long total = oracle_edge_testnum(s, qs, qe, N, offset, testnum, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  edge_job jobs[256];
  for (int t = 0; t < nthreads; ++t) {
    edge_job e = {s, qs, qe, res, testnum, N, offset, t, nthreads};
    jobs[t] = e;
    if (nthreads == 1) edge_worker(&jobs[0]);
    else pthread_create(&th[t], NULL, edge_worker, &jobs[t]);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(testnum);
  return total;
}

/* ---------------- scene construction ---------------- */

void oracle_scene_free(oracle_scene* s) {
  if (!s) return;
  free(s->sys); free(s->n); free(s->m); free(s->dspn); free(s->dspm);
  free(s->vertices); free(s->dsp); free(s->cnt); free(s->dest);
  free(s->pair_dsp); free(s->pair_cnt); free(s->pair_dest);
  free(s);
}

/* polytope4data::build (polytope4data.h:296-366) */
oracle_scene* oracle_scene_create(int ndof, const float* a, const float* alpha,
                                  const float* q, const float* d, const int* types,
                                  const float* mins, const float* maxs, float dq,
                                  int N, const int* sys, const int* n, const int* m,
                                  const float* vertices, const int* dsp, const int* cnt,
                                  const int* dest, int M, const int* from, const int* to) {
  if (ndof < 1 || ndof > ORACLE_MAXDOF || N < 1 || M < 0) return NULL;
  for (int i = 1; i < M; ++i)
    if (from[i] < from[i - 1]) return NULL;
  oracle_scene* s = (oracle_scene*)calloc(1, sizeof(oracle_scene));
  s->ndof = ndof;
  for (int i = 0; i < ndof; ++i) {
    s->a[i] = a[i]; s->alpha[i] = alpha[i]; s->q[i] = q[i]; s->d[i] = d[i];
    s->types[i] = types[i]; s->mins[i] = mins[i]; s->maxs[i] = maxs[i];
  }
  s->dq = dq;
  s->N = N;
  s->sys = (int*)malloc(sizeof(int) * N);
  s->n = (int*)malloc(sizeof(int) * N);
  s->m = (int*)malloc(sizeof(int) * N);
  s->dspn = (int*)malloc(sizeof(int) * N);
  s->dspm = (int*)malloc(sizeof(int) * N);
  int sumn = 0, summ = 0;
  for (int k = 0; k < N; ++k) {
    if (sys[k] < 0 || sys[k] > ndof || n[k] < 1 || m[k] < 0) { s->sumn = 0; oracle_scene_free(s); return NULL; }
    s->sys[k] = sys[k]; s->n[k] = n[k]; s->m[k] = m[k];
    s->dspn[k] = sumn; s->dspm[k] = summ;
    sumn += n[k]; summ += m[k];
  }
  s->sumn = sumn; s->summ = summ;
  s->vertices = (ofloat4*)malloc(sizeof(ofloat4) * sumn);
  s->dsp = (int*)malloc(sizeof(int) * sumn);
  s->cnt = (int*)malloc(sizeof(int) * sumn);
  s->dest = (int*)malloc(sizeof(int) * (summ > 0 ? summ : 1));
  for (int i = 0; i < sumn; ++i) {
    s->vertices[i].x = vertices[3 * i];
    s->vertices[i].y = vertices[3 * i + 1];
    s->vertices[i].z = vertices[3 * i + 2];
    s->vertices[i].w = 1.0f;
    s->dsp[i] = dsp[i];
    s->cnt[i] = cnt[i];
  }
  for (int i = 0; i < summ; ++i) s->dest[i] = dest[i];
  s->M = M;
  s->pair_dsp = (int*)malloc(sizeof(int) * N);
  s->pair_cnt = (int*)calloc((size_t)N, sizeof(int));
  s->pair_dest = (int*)malloc(sizeof(int) * (M > 0 ? M : 1));
  for (int i = 0; i < M; ++i) {
    if (from[i] < 0 || from[i] >= N || to[i] < 0 || to[i] >= N) { oracle_scene_free(s); return NULL; }
    s->pair_dest[i] = to[i];
    s->pair_cnt[from[i]]++;
  }
  int num = 0;
  for (int k = 0; k < N; ++k) { s->pair_dsp[k] = num; num += s->pair_cnt[k]; }
  return s;
}

static int read_raw(const char* dir, const char* name, void* dst, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  size_t got = fread(dst, 1, bytes, f);
  fclose(f);
  return got == bytes ? 0 : -1;
}

/* RobotConfigspace::load_config (robotconfigspace.cpp:100-208) */
oracle_scene* oracle_scene_load_dir(const char* dir, float dq) {
  int ndof = 0;
  if (read_raw(dir, "ndof.bin", &ndof, 4) || ndof < 1 || ndof > ORACLE_MAXDOF) return NULL;
  float a[ORACLE_MAXDOF], al[ORACLE_MAXDOF], q[ORACLE_MAXDOF], d[ORACLE_MAXDOF];
  float mins[ORACLE_MAXDOF], maxs[ORACLE_MAXDOF];
  int types[ORACLE_MAXDOF];
  size_t fb = sizeof(float) * (size_t)ndof;
  if (read_raw(dir, "dh/a.bin", a, fb) || read_raw(dir, "dh/alpha.bin", al, fb) ||
      read_raw(dir, "dh/q.bin", q, fb) || read_raw(dir, "dh/d.bin", d, fb) ||
      read_raw(dir, "dh/types.bin", types, sizeof(int) * (size_t)ndof) ||
      read_raw(dir, "mins.bin", mins, fb) || read_raw(dir, "maxs.bin", maxs, fb))
    return NULL;
  int N = 0, M = 0;
  if (read_raw(dir, "polys/N.bin", &N, 4) || N <= 0) return NULL;
  if (read_raw(dir, "pairs/M.bin", &M, 4) || M <= 0) return NULL;
  int* sys = (int*)malloc(sizeof(int) * N);
  int* from = (int*)malloc(sizeof(int) * M);
  int* to = (int*)malloc(sizeof(int) * M);
  int* n = (int*)malloc(sizeof(int) * N);
  int* m = (int*)malloc(sizeof(int) * N);
  oracle_scene* s = NULL;
  float* verts = NULL; int *dsp = NULL, *cnt = NULL, *dest = NULL;
  if (read_raw(dir, "polys/sys.bin", sys, sizeof(int) * N) ||
      read_raw(dir, "pairs/from.bin", from, sizeof(int) * M) ||
      read_raw(dir, "pairs/to.bin", to, sizeof(int) * M))
    goto done;
  int sumn = 0, summ = 0;
  for (int i = 0; i < N; ++i) {
    char name[256]; int size[2];
    snprintf(name, sizeof name, "polys/poly%d/size.bin", i);
    if (read_raw(dir, name, size, 8)) goto done;
    n[i] = size[0]; m[i] = size[1];
    if (n[i] < 1 || m[i] < 0) goto done;
    sumn += n[i]; summ += m[i];
  }
  verts = (float*)malloc(sizeof(float) * 3 * sumn);
  dsp = (int*)malloc(sizeof(int) * sumn);
  cnt = (int*)malloc(sizeof(int) * sumn);
  dest = (int*)malloc(sizeof(int) * (summ > 0 ? summ : 1));
  int on = 0, om = 0;
  for (int i = 0; i < N; ++i) {
    char name[256];
    snprintf(name, sizeof name, "polys/poly%d/vertices.bin", i);
    if (read_raw(dir, name, verts + 3 * on, sizeof(float) * 3 * n[i])) goto done;
    snprintf(name, sizeof name, "polys/poly%d/dsp.bin", i);
    if (read_raw(dir, name, dsp + on, sizeof(int) * n[i])) goto done;
    snprintf(name, sizeof name, "polys/poly%d/cnt.bin", i);
    if (read_raw(dir, name, cnt + on, sizeof(int) * n[i])) goto done;
    snprintf(name, sizeof name, "polys/poly%d/dest.bin", i);
    if (m[i] > 0 && read_raw(dir, name, dest + om, sizeof(int) * m[i])) goto done;
    on += n[i]; om += m[i];
  }
  s = oracle_scene_create(ndof, a, al, q, d, types, mins, maxs, dq, N, sys, n, m, verts, dsp,
                          cnt, dest, M, from, to);
done:
  free(sys); free(from); free(to); free(n); free(m);
  free(verts); free(dsp); free(cnt); free(dest);
  return s;
}
