// scene.cpp -- scene validation, the reference's directory format, hull
// adjacency and the seeded benchmark scene generators.
#include "scene.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cerrno>
#include <cstring>
#include <set>
#include <sys/stat.h>

namespace gprm {

namespace {

std::string join(const std::string& a, const std::string& b) { return a + "/" + b; }

bool read_bytes(const std::string& path, void* dst, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  size_t got = std::fread(dst, 1, bytes, f);
  std::fclose(f);
  return got == bytes;
}

bool write_bytes(const std::string& path, const void* src, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  size_t put = bytes ? std::fwrite(src, 1, bytes, f) : 0;
  bool ok = std::fclose(f) == 0 && put == bytes;
  return ok;
}

bool mkdir_p(const std::string& d) {
  struct stat st;
  if (stat(d.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
  size_t pos = d.find_last_of('/');
  if (pos != std::string::npos && pos > 0 && !mkdir_p(d.substr(0, pos))) return false;
  return ::mkdir(d.c_str(), 0755) == 0 || errno == EEXIST;
}

}  // namespace

std::string validate(const Scene& s) {
  if (s.ndof < 1 || s.ndof > kMaxDof) return "ndof out of range";
  const int N = (int)s.polys.size();
  if (N < 1) return "no polytopes";
  if ((int)s.sys.size() != N) return "sys size != number of polytopes";
  for (int k = 0; k < N; ++k) {
    const Polytope& p = s.polys[k];
    if (s.sys[k] < 0 || s.sys[k] > s.ndof) return "sys index out of range";
    if (p.n() < 1 || (int)p.verts.size() != 3 * p.n() || (int)p.dsp.size() != p.n())
      return "malformed polytope arrays";
    for (int i = 0; i < p.n(); ++i) {
      // robotconfigspace.cpp:45-49: every vertex needs at least one edge
      if (p.cnt[i] < 1) return "polytope vertex without edges";
      if (p.dsp[i] < 0 || p.dsp[i] + p.cnt[i] > p.m()) return "adjacency row out of range";
    }
    for (int v : p.dest)
      if (v < 0 || v >= p.n()) return "adjacency target out of range";
    for (int i = 0; i < 3 * p.n(); ++i)
      if (!std::isfinite(p.verts[i])) return "non-finite vertex";
  }
  if (s.from.size() != s.to.size() || s.from.empty()) return "empty or mismatched pair list";
  for (size_t i = 0; i < s.from.size(); ++i) {
    if (s.from[i] < 0 || s.from[i] >= N || s.to[i] < 0 || s.to[i] >= N)
      return "pair index out of range";
    if (i > 0 && s.from[i] < s.from[i - 1]) return "pairs not sorted by from";
  }
  for (int i = 0; i < s.ndof; ++i)
    if (s.types[i] != 0 && s.types[i] != 1) return "joint type must be 0 or 1";
  return "";
}

// robotconfigspace.cpp:100-208 and util.hpp:130-149 (raw little-endian)
std::string load_dir(const std::string& dir, Scene* out) {
  Scene s;
  if (!read_bytes(join(dir, "ndof.bin"), &s.ndof, 4)) return "cannot read ndof.bin in " + dir;
  if (s.ndof < 1 || s.ndof > kMaxDof) return "ndof out of range";
  size_t fb = sizeof(float) * s.ndof;
  if (!read_bytes(join(dir, "dh/a.bin"), s.a, fb) || !read_bytes(join(dir, "dh/alpha.bin"), s.alpha, fb) ||
      !read_bytes(join(dir, "dh/q.bin"), s.q, fb) || !read_bytes(join(dir, "dh/d.bin"), s.d, fb) ||
      !read_bytes(join(dir, "dh/types.bin"), s.types, sizeof(int) * s.ndof) ||
      !read_bytes(join(dir, "mins.bin"), s.mins, fb) || !read_bytes(join(dir, "maxs.bin"), s.maxs, fb))
    return "cannot read robot files";
  int N = 0, M = 0;
  if (!read_bytes(join(dir, "polys/N.bin"), &N, 4) || N <= 0 || N > (1 << 20)) return "bad polys/N.bin";
  if (!read_bytes(join(dir, "pairs/M.bin"), &M, 4) || M <= 0 || M > (1 << 24)) return "bad pairs/M.bin";
  s.sys.resize(N);
  s.from.resize(M);
  s.to.resize(M);
  if (!read_bytes(join(dir, "polys/sys.bin"), s.sys.data(), 4 * (size_t)N) ||
      !read_bytes(join(dir, "pairs/from.bin"), s.from.data(), 4 * (size_t)M) ||
      !read_bytes(join(dir, "pairs/to.bin"), s.to.data(), 4 * (size_t)M))
    return "cannot read sys/pairs";
  s.polys.resize(N);
  for (int i = 0; i < N; ++i) {
    std::string pd = dir + "/polys/poly" + std::to_string(i);
    int size[2];
    if (!read_bytes(join(pd, "size.bin"), size, 8)) return "cannot read " + pd + "/size.bin";
    int n = size[0], m = size[1];
    if (n < 1 || n > (1 << 16) || m < 0 || m > (1 << 22)) return "bad polytope size in " + pd;
    Polytope& p = s.polys[i];
    p.verts.resize(3 * (size_t)n);
    p.dsp.resize(n);
    p.cnt.resize(n);
    p.dest.resize(m);
    if (!read_bytes(join(pd, "vertices.bin"), p.verts.data(), 12 * (size_t)n) ||
        !read_bytes(join(pd, "dsp.bin"), p.dsp.data(), 4 * (size_t)n) ||
        !read_bytes(join(pd, "cnt.bin"), p.cnt.data(), 4 * (size_t)n) ||
        (m > 0 && !read_bytes(join(pd, "dest.bin"), p.dest.data(), 4 * (size_t)m)))
      return "cannot read polytope files in " + pd;
  }
  std::string err = validate(s);
  if (!err.empty()) return err;
  *out = std::move(s);
  return "";
}

std::string write_dir(const Scene& s, const std::string& dir) {
  std::string err = validate(s);
  if (!err.empty()) return err;
  if (!mkdir_p(dir + "/dh") || !mkdir_p(dir + "/polys") || !mkdir_p(dir + "/pairs"))
    return "cannot create directories under " + dir;
  size_t fb = sizeof(float) * s.ndof;
  int N = (int)s.polys.size(), M = (int)s.from.size();
  bool ok = write_bytes(join(dir, "ndof.bin"), &s.ndof, 4) && write_bytes(join(dir, "dh/a.bin"), s.a, fb) &&
            write_bytes(join(dir, "dh/alpha.bin"), s.alpha, fb) && write_bytes(join(dir, "dh/q.bin"), s.q, fb) &&
            write_bytes(join(dir, "dh/d.bin"), s.d, fb) &&
            write_bytes(join(dir, "dh/types.bin"), s.types, sizeof(int) * s.ndof) &&
            write_bytes(join(dir, "mins.bin"), s.mins, fb) && write_bytes(join(dir, "maxs.bin"), s.maxs, fb) &&
            write_bytes(join(dir, "polys/N.bin"), &N, 4) &&
            write_bytes(join(dir, "polys/sys.bin"), s.sys.data(), 4 * (size_t)N) &&
            write_bytes(join(dir, "pairs/M.bin"), &M, 4) &&
            write_bytes(join(dir, "pairs/from.bin"), s.from.data(), 4 * (size_t)M) &&
            write_bytes(join(dir, "pairs/to.bin"), s.to.data(), 4 * (size_t)M);
  if (!ok) return "cannot write robot/pair files under " + dir;
  for (int i = 0; i < N; ++i) {
    const Polytope& p = s.polys[i];
    std::string pd = dir + "/polys/poly" + std::to_string(i);
    if (!mkdir_p(pd)) return "cannot create " + pd;
    int size[2] = {p.n(), p.m()};
    ok = write_bytes(join(pd, "size.bin"), size, 8) &&
         write_bytes(join(pd, "vertices.bin"), p.verts.data(), 12 * (size_t)p.n()) &&
         write_bytes(join(pd, "dsp.bin"), p.dsp.data(), 4 * (size_t)p.n()) &&
         write_bytes(join(pd, "cnt.bin"), p.cnt.data(), 4 * (size_t)p.n()) &&
         write_bytes(join(pd, "dest.bin"), p.dest.data(), 4 * (size_t)p.m());
    if (!ok) return "cannot write " + pd;
  }
  return "";
}

// ---------------------------------------------------------------------------
// Hull adjacency.
//
// For the small point sets of the workloads (<= 64 points) a face search over
// all triples is exact enough in double precision and simple: a triple is a
// hull facet when every other point lies strictly on one side of its plane.
// Degenerate sets (4 coplanar points, or any point within a relative 1e-12 of
// a facet plane) are rejected so generators can resample.
bool hull_polytope(const std::vector<double>& pts, Polytope* out) {
  const int n = (int)pts.size() / 3;
  if (n < 4 || n > 64) return false;
  auto P = [&](int i, int c) { return pts[3 * i + c]; };
  double scale = 0.0;
  for (double v : pts) scale = std::max(scale, std::fabs(v));
  std::vector<std::set<int>> nb(n);
  int faces = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      for (int k = j + 1; k < n; ++k) {
        double ux = P(j, 0) - P(i, 0), uy = P(j, 1) - P(i, 1), uz = P(j, 2) - P(i, 2);
        double vx = P(k, 0) - P(i, 0), vy = P(k, 1) - P(i, 1), vz = P(k, 2) - P(i, 2);
        double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
        double nn = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (nn < 1e-14 * scale * scale) return false;  // collinear triple
        int pos = 0, neg = 0;
        for (int l = 0; l < n; ++l) {
          if (l == i || l == j || l == k) continue;
          double s = (P(l, 0) - P(i, 0)) * nx + (P(l, 1) - P(i, 1)) * ny + (P(l, 2) - P(i, 2)) * nz;
          if (std::fabs(s) < 1e-12 * nn * scale) return false;  // coplanar: degenerate
          if (s > 0) ++pos; else ++neg;
          if (pos && neg) break;
        }
        if (pos && neg) continue;
        ++faces;
        nb[i].insert(j); nb[i].insert(k);
        nb[j].insert(i); nb[j].insert(k);
        nb[k].insert(i); nb[k].insert(j);
      }
  std::vector<int> keep;
  for (int i = 0; i < n; ++i)
    if (!nb[i].empty()) keep.push_back(i);
  const int V = (int)keep.size();
  if (V < 4 || faces != 2 * V - 4) return false;  // must be a triangulated sphere
  std::vector<int> remap(n, -1);
  for (int i = 0; i < V; ++i) remap[keep[i]] = i;
  Polytope p;
  size_t E2 = 0;
  for (int i : keep) E2 += nb[i].size();
  if (E2 != (size_t)(6 * V - 12)) return false;
  for (int i : keep) {
    p.verts.push_back((float)P(i, 0));
    p.verts.push_back((float)P(i, 1));
    p.verts.push_back((float)P(i, 2));
    p.dsp.push_back((int)p.dest.size());
    p.cnt.push_back((int)nb[i].size());
    std::vector<int> row;
    for (int j : nb[i]) row.push_back(remap[j]);
    std::sort(row.begin(), row.end());
    for (int j : row) p.dest.push_back(j);
  }
  *out = std::move(p);
  return true;
}

namespace {

// Box with a fixed triangulation of its faces: vertex index 4*ix+2*iy+iz,
// 12 box edges plus one diagonal per face (rows ascending, symmetric).
Polytope box_polytope(const double corner[8][3]) {
  static const int edges[18][2] = {
      {0, 1}, {0, 2}, {0, 4}, {1, 3}, {1, 5}, {2, 3}, {2, 6}, {3, 7}, {4, 5}, {4, 6}, {5, 7}, {6, 7},
      {0, 3}, {4, 7}, {0, 5}, {2, 7}, {0, 6}, {1, 7}};
  std::vector<std::set<int>> nb(8);
  for (auto& e : edges) {
    nb[e[0]].insert(e[1]);
    nb[e[1]].insert(e[0]);
  }
  Polytope p;
  for (int i = 0; i < 8; ++i) {
    for (int c = 0; c < 3; ++c) p.verts.push_back((float)corner[i][c]);
    p.dsp.push_back((int)p.dest.size());
    p.cnt.push_back((int)nb[i].size());
    for (int j : nb[i]) p.dest.push_back(j);
  }
  return p;
}

Polytope link_box(double L, double w) {
  double c[8][3];
  for (int i = 0; i < 8; ++i) {
    c[i][0] = (i & 4) ? L : 0.0;
    c[i][1] = (i & 2) ? 0.5 * w : -0.5 * w;
    c[i][2] = (i & 1) ? 0.5 * w : -0.5 * w;
  }
  return box_polytope(c);
}

// Uniform rotation: uniform point in the unit 4-ball by rejection, normalised
// to a unit quaternion (arithmetic and sqrt only, hence identical everywhere).
void random_rotation(Rng& r, double R[3][3]) {
  double w, x, y, z, n2;
  do {
    w = r.uniform(-1, 1); x = r.uniform(-1, 1); y = r.uniform(-1, 1); z = r.uniform(-1, 1);
    n2 = w * w + x * x + y * y + z * z;
  } while (n2 > 1.0 || n2 < 1e-6);
  double n = std::sqrt(n2);
  w /= n; x /= n; y /= n; z /= n;
  R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z);     R[0][2] = 2 * (x * z + w * y);
  R[1][0] = 2 * (x * y + w * z);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
  R[2][0] = 2 * (x * z - w * y);     R[2][1] = 2 * (y * z + w * x);     R[2][2] = 1 - 2 * (x * x + y * y);
}

void random_center(Rng& r, double c[3]) {
  do {
    for (int i = 0; i < 3; ++i) c[i] = r.uniform(-3.0, 3.0);
  } while (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] < 0.25);
}

Polytope obstacle_box(Rng& r) {
  double c[3], h[3], R[3][3], corner[8][3];
  random_center(r, c);
  for (int i = 0; i < 3; ++i) h[i] = r.uniform(0.2, 0.6);
  random_rotation(r, R);
  for (int i = 0; i < 8; ++i) {
    double l[3] = {(i & 4) ? h[0] : -h[0], (i & 2) ? h[1] : -h[1], (i & 1) ? h[2] : -h[2]};
    for (int a = 0; a < 3; ++a) corner[i][a] = c[a] + R[a][0] * l[0] + R[a][1] * l[1] + R[a][2] * l[2];
  }
  return box_polytope(corner);
}

Polytope obstacle_hull(Rng& r, int npts) {
  for (;;) {
    double c[3];
    random_center(r, c);
    double rad = r.uniform(0.2, 0.5);
    std::vector<double> pts;
    for (int i = 0; i < npts; ++i) {
      double x, y, z;
      do {
        x = r.uniform(-1, 1); y = r.uniform(-1, 1); z = r.uniform(-1, 1);
      } while (x * x + y * y + z * z > 1.0);
      pts.push_back(c[0] + rad * x);
      pts.push_back(c[1] + rad * y);
      pts.push_back(c[2] + rad * z);
    }
    Polytope p;
    if (hull_polytope(pts, &p)) return p;
  }
}

// Hull of npts points on a sphere of radius in [rmin, rmax] (every point a
// vertex): many vertices per obstacle.
Polytope obstacle_shell(Rng& r, int npts, double rmin, double rmax) {
  for (;;) {
    double c[3];
    random_center(r, c);
    double rad = r.uniform(rmin, rmax);
    std::vector<double> pts;
    for (int i = 0; i < npts; ++i) {
      double x, y, z, n2;
      do {
        x = r.uniform(-1, 1); y = r.uniform(-1, 1); z = r.uniform(-1, 1);
        n2 = x * x + y * y + z * z;
      } while (n2 > 1.0 || n2 < 0.05);
      const double inv = 1.0 / std::sqrt(n2);
      pts.push_back(c[0] + rad * x * inv);
      pts.push_back(c[1] + rad * y * inv);
      pts.push_back(c[2] + rad * z * inv);
    }
    Polytope p;
    if (hull_polytope(pts, &p)) return p;
  }
}

void set_arm(Scene* s, int ndof, double a, const double* alpha_cycle, int ncycle, double link_w,
             int nobs_polys) {
  s->ndof = ndof;
  const float fpi = (float)M_PI;
  for (int i = 0; i < ndof; ++i) {
    s->a[i] = (float)a;
    s->alpha[i] = (float)alpha_cycle[i % ncycle];
    s->q[i] = 0.0f;
    s->d[i] = 0.0f;
    s->types[i] = 0;
    s->mins[i] = -fpi;
    s->maxs[i] = fpi;
  }
  // Link bodies: frame s (1..ndof) carries the segment to the next joint,
  // x in [0, a], square cross-section link_w.
  for (int l = 1; l <= ndof; ++l) {
    s->polys.push_back(link_box(a, link_w));
    s->sys.push_back(l);
  }
  // Pairs: every obstacle against every link, then non-adjacent link pairs.
  for (int e = 0; e < nobs_polys; ++e)
    for (int l = 0; l < ndof; ++l) {
      s->from.push_back(e);
      s->to.push_back(nobs_polys + l);
    }
  for (int i = 0; i < ndof; ++i)
    for (int j = i + 2; j < ndof; ++j) {
      s->from.push_back(nobs_polys + i);
      s->to.push_back(nobs_polys + j);
    }
}

}  // namespace

std::string generate(int workload, uint64_t seed, Scene* out) {
  Scene s;
  Rng r(seed * 0x2545F4914F6CDD1DULL + (uint64_t)workload);
  const double hp = M_PI / 2;
  std::vector<Polytope> obs;
  int ndof = 0;
  double a = 0, w = 0;
  std::vector<double> cyc;
  if (workload == 1) {
    for (int i = 0; i < 8; ++i) obs.push_back(obstacle_box(r));
    ndof = 4; a = 1.0; w = 0.2; cyc = {0.0, hp};
  } else if (workload == 2) {
    for (int i = 0; i < 16; ++i) obs.push_back(obstacle_hull(r, 12));
    ndof = 6; a = 0.8; w = 0.15; cyc = {0.0, hp, -hp};
  } else if (workload == 3) {
    for (int i = 0; i < 32; ++i) obs.push_back(obstacle_hull(r, 16));
    ndof = 7; a = 0.6; w = 0.12; cyc = {0.0, hp, -hp};
  } else if (workload == 6) {
    // a scene whose image exceeds LDS (about 300 KB): 256 small obstacles of
    // 32 vertices each
    for (int i = 0; i < 256; ++i) obs.push_back(obstacle_shell(r, 32, 0.06, 0.15));
    ndof = 7; a = 0.6; w = 0.12; cyc = {0.0, hp, -hp};
  } else if (workload == 5) {
    int nobs = r.integer(8, 128);
    for (int i = 0; i < nobs; ++i) obs.push_back(obstacle_hull(r, r.integer(8, 24)));
    ndof = 7; a = 0.6; w = 0.12; cyc = {0.0, hp, -hp};
  } else {
    return "unknown workload id";
  }
  // obstacles first (world frame, sys 0), then the links
  for (auto& p : obs) {
    s.polys.push_back(p);
    s.sys.push_back(0);
  }
  std::vector<Polytope> tmp;
  set_arm(&s, ndof, a, cyc.data(), (int)cyc.size(), w, (int)obs.size());
  std::string err = validate(s);
  if (!err.empty()) return err;
  *out = std::move(s);
  return "";
}

}  // namespace gprm
