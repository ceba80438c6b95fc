"""The C++ drop-in (include/gprm_configspace.hpp) on the GPU: a program that
only sees Configspace<ndof>* (as the reference's PRMSolver does,
prm/src/lib/prm/prmsolver.h:390) calls indicator(q), the batched indicator,
indicator2, two indicator2_async requests in flight with their waits, and
the compressed indicator2 (configspace.hpp:16-47); the answers are compared
with the golden vectors of the reference's own CPU path and with the oracle.
The program also times single-configuration and small edge calls."""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include "gprm_configspace.hpp"
#include <chrono>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

template <class T> std::vector<T> rd(const std::string& p, size_t n) {
  std::vector<T> v(n);
  std::ifstream f(p, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(sizeof(T) * n));
  return v;
}
template <class T> void wr(const std::string& p, const std::vector<T>& v) {
  std::ofstream f(p, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(sizeof(T) * v.size()));
}
static double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// everything below sees only the interface, as a planner does
static int run(Configspace<NDOF>* space, const std::string& w, int NQ, int NE) {
  const std::vector<float> Q = rd<float>(w + "/Q.bin", (size_t)NDOF * NQ);
  const std::vector<float> QS = rd<float>(w + "/QS.bin", (size_t)NDOF * NE);
  const std::vector<float> QE = rd<float>(w + "/QE.bin", (size_t)NDOF * NE);
  std::vector<int> c1(NQ), cb(NQ), e2(NE), ea(NE), ec(2 * (size_t)NE);
  for (int k = 0; k < NQ; ++k) {
    float q[NDOF];
    for (int i = 0; i < NDOF; ++i) q[i] = Q[k + (size_t)NQ * i];
    c1[k] = space->indicator(q);
  }
  if (space->indicator(Q.data(), cb.data(), NQ, NQ) != 0) return 3;
  if (space->indicator2(QS.data(), QE.data(), e2.data(), NE, NE) != 0) return 4;
  // two requests in flight, the halves of the edge set, waited in reverse order
  const int h = NE / 2;
  int r1 = -1, r2 = -1;
  if (space->indicator2_async(QS.data(), QE.data(), ea.data(), h, NE, r1) != 0) return 5;
  if (space->indicator2_async(QS.data() + h, QE.data() + h, ea.data() + h, NE - h, NE, r2) != 0) return 5;
  if (space->indicator2_async_wait(r2) != 0 || space->indicator2_async_wait(r1) != 0) return 6;
  // compressed: start i joins ends 2i (QE[i]) and 2i+1 (QE[i+1]); one offset
  // (2 NE) for both arrays
  const int M = NE, off = 2 * NE;
  std::vector<float> qs((size_t)NDOF * off, 0.0f), qe((size_t)NDOF * off);
  std::vector<int> posqe(M), numqe(M, 2);
  for (int i = 0; i < M; ++i) {
    posqe[i] = 2 * i;
    for (int d = 0; d < NDOF; ++d) {
      qs[i + (size_t)off * d] = QS[i + (size_t)NE * d];
      qe[2 * i + (size_t)off * d] = QE[i + (size_t)NE * d];
      qe[2 * i + 1 + (size_t)off * d] = QE[(i + 1) % NE + (size_t)NE * d];
    }
  }
  if (space->indicator2(qs.data(), M, qe.data(), ec.data(), posqe.data(), numqe.data(), off) != 0) return 7;
  wr(w + "/c1.bin", c1);
  wr(w + "/cb.bin", cb);
  wr(w + "/e2.bin", e2);
  wr(w + "/ea.bin", ea);
  wr(w + "/ec.bin", ec);
  // latency of the calls a reference-style planner makes
  float q[NDOF];
  for (int i = 0; i < NDOF; ++i) q[i] = Q[(size_t)NQ * i];
  for (int r = 0; r < 20; ++r) space->indicator(q);
  double t0 = now_us();
  for (int r = 0; r < 200; ++r) space->indicator(q);
  const double us1 = (now_us() - t0) / 200;
  std::vector<int> e8(8);
  for (int r = 0; r < 5; ++r) space->indicator2(QS.data(), QE.data(), e8.data(), 8, NE);
  t0 = now_us();
  for (int r = 0; r < 100; ++r) space->indicator2(QS.data(), QE.data(), e8.data(), 8, NE);
  const double us8 = (now_us() - t0) / 100;
  std::printf("us_per_indicator1 %.2f\nus_per_indicator2_8_edges %.2f\n", us1, us8);
  return 0;
}

int main(int argc, char** argv) {
  RobotConfigspaceGPU<NDOF> cs(argv[1], std::stof(argv[5]), 1024);
  if (cs.init(0, 1) != 0) return 2;
  return run(&cs, argv[2], std::stoi(argv[3]), std::stoi(argv[4]));
}
'''


@pytest.mark.parametrize("name", ["config1", "wl2"])
def test_cpp_configspace_on_gpu(golden_dir, tmp_path, name):
    import gpuprm_amd
    z, d = golden_dir(name)
    nd = int(z["scene_ndof"])
    Q, QS, QE = z["Q"], z["QS"], z["QE"]
    NQ, NE = Q.shape[1], QS.shape[1]
    w = str(tmp_path)
    for k, a in (("Q", Q), ("QS", QS), ("QE", QE)):
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(w, k + ".bin"))
    src = tmp_path / "dropin.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "dropin"
    libdir = os.path.dirname(gpuprm_amd.LIB_PATH)
    subprocess.run(["g++", "-O2", "-std=c++11", "-Wall", "-Werror", "-DNDOF=%d" % nd, "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", libdir, "-lgprm",
                    "-Wl,-rpath," + libdir], check=True)
    print("drop-in program: %s, %d configurations, %d edges" % (name, NQ, NE), flush=True)
    out = subprocess.run([str(exe), d, w, str(NQ), str(NE), repr(float(z["dq"]))], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.returncode
    rd = lambda k, n: np.fromfile(os.path.join(w, k + ".bin"), np.int32, n)
    assert np.array_equal(rd("c1", NQ), z["codes"])
    assert np.array_equal(rd("cb", NQ), z["codes"])
    want = z["edge_res"] != 0
    assert np.array_equal(rd("e2", NE) != 0, want)
    assert np.array_equal(rd("ea", NE) != 0, want)
    ec = rd("ec", 2 * NE)
    assert np.array_equal(ec[0::2] != 0, want)
    o = OracleScene(d, float(z["dq"]))
    nxt, _ = o.indicator2(QS, np.roll(QE, -1, axis=1).copy(), nthreads=8)
    assert np.array_equal(ec[1::2] != 0, nxt != 0)
    lat = dict(l.split() for l in out.stdout.splitlines() if l.startswith("us_"))
    assert float(lat["us_per_indicator1"]) > 0
