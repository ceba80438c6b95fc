// gprm_configspace.hpp -- header-only C++ drop-in for GPUPRM's
// configuration-space interface, forwarding to the C ABI of gprm.h.
//
// Configspace<ndof> below has the pure virtuals of the reference's interface
// (prm/src/lib/prm/configspace.hpp:11-50); RobotConfigspaceGPU<ndof> plays
// the role of RobotConfigspace<ndof> (prm/src/lib/robotspace/
// robotconfigspace.h:26-176) with the same constructors and methods, so a
// planner written against Configspace<ndof>* (such as the reference's
// PRMSolver) links unchanged.  If the reference's own configspace.hpp is
// already included, define GPRM_HAVE_CONFIGSPACE to reuse its declaration.
#ifndef GPRM_CONFIGSPACE_HPP
#define GPRM_CONFIGSPACE_HPP

#include <stdexcept>
#include <string>

#include "gprm.h"

#ifndef GPRM_HAVE_CONFIGSPACE
template <int ndof>
class Configspace {
 public:
  virtual ~Configspace() {}
  virtual int init(const int ressource_rank = 0, const int ressource_size = 1) = 0;
  virtual int indicator(const float* q, int* res, const int N, const int offset) = 0;
  virtual int indicator(const float* q) = 0;
  virtual int indicator2(const float* qs, const float* qe, int* res, const int N, const int offset) = 0;
  virtual int indicator2_async(const float* qs, const float* qe, int* res, const int N, const int offset,
                               int& request) = 0;
  virtual int indicator2_async_wait(int request) = 0;
  virtual int indicator2(const float* qs, const int M, const float* qe, int* res, const int* posqe,
                         const int* numqe, const int offset) = 0;
  virtual int check_boundaries(const float* q, int* res, const int N, const int offset) = 0;
  virtual int check_boundaries(const float* q) = 0;
  virtual float min(int i) = 0;
  virtual float max(int i) = 0;
  virtual float deltaq() = 0;
  virtual int dim() = 0;
};
#endif

// Thrown where the reference would abort (scene that cannot be loaded).
struct GprmException : std::runtime_error {
  int code;
  GprmException(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <int ndof>
class RobotConfigspaceGPU : public Configspace<ndof> {
 public:
  // RobotConfigspace(std::string path, float dq, int nbuf) (robotconfigspace.h:42)
  RobotConfigspaceGPU(const std::string& path, const float dq, const int nbuf) {
    check_create(gprm_create_from_dir(path.c_str(), dq, nbuf, &ctx_));
  }
  // RobotConfigspace(robot, polys, sys, N, from, to, M, mins, maxs, dq, nbuf)
  // (robotconfigspace.h:34-41), with the robot and polytopes flattened.
  RobotConfigspaceGPU(const float* a, const float* alpha, const float* q, const float* d, const int* types,
                      int N, const int* sys, const int* n, const int* m, const float* vertices, const int* dsp,
                      const int* cnt, const int* dest, const int* from, const int* to, int M, const float* mins,
                      const float* maxs, const float dq, const int nbuf) {
    check_create(gprm_create(ndof, a, alpha, q, d, types, N, sys, n, m, vertices, dsp, cnt, dest, from, to, M, mins,
                             maxs, dq, nbuf, &ctx_));
  }
  // Not marked override: the reference's Configspace<ndof> declares no
  // virtual destructor (configspace.hpp:11-50), and this class must compile
  // against that declaration too.  With the declaration below it is virtual.
  ~RobotConfigspaceGPU() { gprm_destroy(ctx_); }
  RobotConfigspaceGPU(const RobotConfigspaceGPU&) = delete;
  RobotConfigspaceGPU& operator=(const RobotConfigspaceGPU&) = delete;

  int init(const int ressource_rank = 0, const int ressource_size = 1) override {
    return gprm_init(ctx_, ressource_rank, ressource_size);
  }
  int indicator(const float* q, int* res, const int N, const int offset) override {
    return gprm_indicator(ctx_, q, res, N, offset);
  }
  // Returns the reference's code (0 free, 1 out of bounds, l+1 or -1 blocked);
  // a runtime error is reported as blocked (-1), the reference's safe side.
  int indicator(const float* q) override {
    int code = 0;
    return gprm_indicator1(ctx_, q, &code) == GPRM_OK ? code : -1;
  }
  int indicator2(const float* qs, const float* qe, int* res, const int N, const int offset) override {
    return gprm_indicator2(ctx_, qs, qe, res, N, offset);
  }
  int indicator2_async(const float* qs, const float* qe, int* res, const int N, const int offset,
                       int& request) override {
    return gprm_indicator2_async(ctx_, qs, qe, res, N, offset, &request);
  }
  int indicator2_async_wait(int request) override { return gprm_indicator2_wait(ctx_, request); }
  int indicator2(const float* qs, const int M, const float* qe, int* res, const int* posqe, const int* numqe,
                 const int offset) override {
    return gprm_indicator2_compressed(ctx_, qs, M, qe, res, posqe, numqe, offset);
  }
  int check_boundaries(const float* q, int* res, const int N, const int offset) override {
    return gprm_check_boundaries(ctx_, q, res, N, offset);
  }
  int check_boundaries(const float* q) override {
    int out = 1;
    return gprm_check_boundaries1(ctx_, q, &out) == GPRM_OK ? out : 1;
  }
  float min(int i) override { return gprm_min(ctx_, i); }
  float max(int i) override { return gprm_max(ctx_, i); }
  float deltaq() override { return gprm_deltaq(ctx_); }
  int dim() override { return ndof; }
  int get_numthreads_all() const { return (int)gprm_numthreads_all(ctx_); }
  gprm_ctx* handle() const { return ctx_; }
  const char* last_error() const { return gprm_last_error(ctx_); }

 private:
  void check_create(int rc) {
    if (rc != GPRM_OK) throw GprmException(rc, gprm_last_error(nullptr));
    if (gprm_dim(ctx_) != ndof) {
      gprm_destroy(ctx_);
      ctx_ = nullptr;
      throw GprmException(GPRM_E_SCENE, "scene has a different number of joints than the template argument");
    }
  }
  gprm_ctx* ctx_ = nullptr;
};

#endif  // GPRM_CONFIGSPACE_HPP
