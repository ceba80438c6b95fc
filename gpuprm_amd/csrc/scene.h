// scene.h -- robot + polytope scene: in-memory form, the reference's on-disk
// directory format, and seeded generators for the benchmark workloads.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gprm {

constexpr int kMaxDof = 16;

// One convex polytope: vertices plus hull adjacency in CRS form, as the
// reference's `polytope` (prm/src/lib/collision4/polytope.h:5-16) with float
// vertices (the on-disk format is float, robotconfigspace.cpp:180-187).
struct Polytope {
  std::vector<float> verts;  // 3*n
  std::vector<int> dsp, cnt; // n, local offsets into dest
  std::vector<int> dest;     // m, local vertex ids
  int n() const { return (int)cnt.size(); }
  int m() const { return (int)dest.size(); }
};

// Scene = robot (modified DH, prm/src/lib/robotspace/robot.h:10-22), bounds,
// polytopes with their frame index sys[k] (0 = world, i = link i), and the
// collision pair list (from, to) sorted by `from`.
struct Scene {
  int ndof = 0;
  float a[kMaxDof] = {}, alpha[kMaxDof] = {}, q[kMaxDof] = {}, d[kMaxDof] = {};
  int types[kMaxDof] = {};
  float mins[kMaxDof] = {}, maxs[kMaxDof] = {};
  std::vector<Polytope> polys;
  std::vector<int> sys;
  std::vector<int> from, to;
};

// Returns "" on success, else an error message.
std::string validate(const Scene& s);
// Reference directory format (robotconfigspace.cpp:100-208).
std::string load_dir(const std::string& dir, Scene* out);
std::string write_dir(const Scene& s, const std::string& dir);

// Seeded generators.  Workload ids:
//   1: 4-DOF, a=1, alpha 0,+pi/2,..., box links 1.0x0.2x0.2, 8 obstacle boxes
//   2: 6-DOF, a=0.8, alpha 0,+pi/2,-pi/2,..., links 0.8x0.15x0.15, 16 hulls of 12 points
//   3: 7-DOF, a=0.6, same alpha cycle, links 0.6x0.12x0.12, 32 hulls of 16 points
//   5: 7-DOF arm of workload 3, obstacle count U{8..128}, hull points U{8..24}
std::string generate(int workload, uint64_t seed, Scene* out);

// Triangulated convex hull of points (double precision), adjacency in the
// convention of the reference's scene tooling (prm/matlab/lib/polygen2.m):
// interior points dropped and re-indexed, symmetric, every row ascending.
// Returns false if the points are degenerate.
bool hull_polytope(const std::vector<double>& pts, Polytope* out);

// SplitMix64 (a fixed, platform-independent stream).
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  int integer(int lo, int hi) { return lo + (int)(uniform() * (double)(hi - lo + 1)); }  // [lo,hi]
};

}  // namespace gprm
