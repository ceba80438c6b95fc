// ref_harness.cpp -- exposes the reference's own CPU collision path through a
// small C ABI so tests can compare against it and bench.py can time it.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from the
// reference sources where they lie (REF=<reference>/prm/src/lib), output only
// into oracle/_ref/.  Nothing here is shipped or called by the product.
//
// The reference TU prm/src/lib/robotspace/robotconfigspace.cpp is compiled
// without GPU_VERSION, i.e. its host path: indicator(q) (:312-346) and the
// host loop of kernel_indicator2 behind indicator2 (:389-465,694-705).  The
// reference instantiates only RobotConfigspace<4> (:809); this TU adds the
// other arm sizes.  Members are opened up (`private` -> `public`) only to read
// the forward-kinematics frames and run single pairs for golden vectors.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#define private public
#include "robotspace/robotconfigspace.cpp"
#undef private

template class RobotConfigspace<2>;
template class RobotConfigspace<3>;
template class RobotConfigspace<5>;
template class RobotConfigspace<6>;
template class RobotConfigspace<7>;
template class RobotConfigspace<8>;

namespace {

struct Base {
  virtual ~Base() {}
  virtual int ndof() = 0;
  virtual int indicator1(const float* q) = 0;
  virtual int indicator2(const float* qs, const float* qe, int* res, int N, int offset) = 0;
  virtual void fk(const float* q, float* out) = 0;
  virtual int sva(const float* q, int k0, int k1, int* iterations) = 0;
  virtual long long numthreads_all() = 0;
};

template <int D>
struct Impl : Base {
  RobotConfigspace<D>* cs;
  explicit Impl(RobotConfigspace<D>* c) : cs(c) {}
  ~Impl() { delete cs; }
  int ndof() override { return D; }
  int indicator1(const float* q) override { return cs->indicator(q); }
  int indicator2(const float* qs, const float* qe, int* res, int N, int offset) override {
    return cs->indicator2(qs, qe, res, N, offset);
  }
  void fk(const float* q, float* out) override {
    cs->kin->calculate(q, 1);
    for (int i = 0; i <= D; ++i) {
      const geo4::trafo4& t = cs->kin->trafos[i];
      for (int c = 0; c < 4; ++c) {
        out[i * 12 + c * 3 + 0] = t.col[c].x;
        out[i * 12 + c * 3 + 1] = t.col[c].y;
        out[i * 12 + c * 3 + 2] = t.col[c].z;
      }
    }
  }
  int sva(const float* q, int k0, int k1, int* iterations) override {
    cs->kin->calculate(q, 1);
    collision4::polytope4 p0, p1;
    cs->polydata->get_polytope(p0, k0);
    cs->polydata->get_polytope(p1, k1);
    return collision4::seperating_vector_algorithm(
        p0, p1, cs->kin->trafos[cs->polydata->sys[k0]], cs->kin->trafos[cs->polydata->sys[k1]],
        iterations);
  }
  long long numthreads_all() override { return cs->get_numthreads_all(); }
};

template <int D>
Base* make(const char* path, float dq, int nbuf) {
  // The reference prints scene statistics on construction/init; keep stdout clean.
  std::streambuf* old = std::cout.rdbuf();
  std::ostringstream sink;
  std::cout.rdbuf(sink.rdbuf());
  RobotConfigspace<D>* c = new RobotConfigspace<D>(std::string(path), dq, nbuf);
  c->init(0, 1);
  std::cout.rdbuf(old);
  return new Impl<D>(c);
}

}  // namespace

extern "C" {

void* ref_create_dir(const char* path, float dq, int nbuf) {
  int nd = 0;
  std::string p = std::string(path) + "/ndof.bin";
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return nullptr;
  if (std::fread(&nd, 4, 1, f) != 1) nd = 0;
  std::fclose(f);
  switch (nd) {
    case 2: return make<2>(path, dq, nbuf);
    case 3: return make<3>(path, dq, nbuf);
    case 4: return make<4>(path, dq, nbuf);
    case 5: return make<5>(path, dq, nbuf);
    case 6: return make<6>(path, dq, nbuf);
    case 7: return make<7>(path, dq, nbuf);
    case 8: return make<8>(path, dq, nbuf);
    default: return nullptr;
  }
}
void ref_destroy(void* h) { delete static_cast<Base*>(h); }
int ref_ndof(void* h) { return static_cast<Base*>(h)->ndof(); }
int ref_indicator1(void* h, const float* q) { return static_cast<Base*>(h)->indicator1(q); }
int ref_indicator2(void* h, const float* qs, const float* qe, int* res, int N, int offset) {
  return static_cast<Base*>(h)->indicator2(qs, qe, res, N, offset);
}
void ref_fk(void* h, const float* q, float* out) { static_cast<Base*>(h)->fk(q, out); }
int ref_sva(void* h, const float* q, int k0, int k1, int* iterations) {
  return static_cast<Base*>(h)->sva(q, k0, k1, iterations);
}
long long ref_numthreads_all(void* h) { return static_cast<Base*>(h)->numthreads_all(); }

// indicator(q) over N SoA items (component i of item k at q[k+offset*i]),
// statically chunked over nthreads handles (one RobotConfigspace per thread:
// the reference object keeps its kinematics frames as a member).
int ref_indicator_mt(void** hs, int nthreads, const float* q, int* res, int N, int offset) {
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([=]() {
      Base* b = static_cast<Base*>(hs[t]);
      const int nd = b->ndof();
      long long lo = (long long)N * t / nthreads, hi = (long long)N * (t + 1) / nthreads;
      float qq[16];
      for (long long k = lo; k < hi; ++k) {
        for (int i = 0; i < nd; ++i) qq[i] = q[k + (long long)offset * i];
        res[k] = b->indicator1(qq);
      }
    });
  }
  for (auto& x : th) x.join();
  return 0;
}

// indicator2 (host loop) over N edges, statically chunked over nthreads handles.
int ref_indicator2_mt(void** hs, int nthreads, const float* qs, const float* qe, int* res, int N, int offset) {
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([=]() {
      Base* b = static_cast<Base*>(hs[t]);
      long long lo = (long long)N * t / nthreads, hi = (long long)N * (t + 1) / nthreads;
      if (hi > lo) b->indicator2(qs + lo, qe + lo, res + lo, (int)(hi - lo), offset);
    });
  }
  for (auto& x : th) x.join();
  return 0;
}

}  // extern "C"
