"""C ABI of libgprm.so: every function declared in include/gprm.h is exported,
and the error paths return GPRM_E_* codes instead of aborting (no device
work: these run on the CPU)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gprm.h")

GPRM_E_INVALID, GPRM_E_SCENE, GPRM_E_HIP, GPRM_E_CAPACITY, GPRM_E_REQUEST, GPRM_E_STATE = -1, -2, -3, -4, -5, -6


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gprm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_interface():
    names = declared()
    for n in ("gprm_create_from_dir", "gprm_create", "gprm_init", "gprm_destroy", "gprm_indicator1",
              "gprm_indicator", "gprm_indicator2", "gprm_indicator2_async", "gprm_indicator2_wait",
              "gprm_indicator2_compressed", "gprm_check_boundaries", "gprm_check_boundaries1", "gprm_min",
              "gprm_max", "gprm_deltaq", "gprm_dim", "gprm_numthreads_all", "gprm_indicator_dev",
              "gprm_indicator2_dev", "gprm_last_error", "gprm_scene_generate", "gprm_profile_enable",
              "gprm_profile_read"):
        assert n in names, n


def test_library_exports_every_declared_symbol():
    import gpuprm_amd
    out = subprocess.run(["nm", "-D", "--defined-only", gpuprm_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(gprm_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    L = gpuprm_amd.lib()
    for n in declared():
        assert hasattr(L, n)


def test_c_header_compiles_as_c(tmp_path):
    src = tmp_path / "use.c"
    src.write_text('#include "gprm.h"\nint main(void){ gprm_ctx* c = 0; (void)c; return GPRM_OK; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                    "-o", str(tmp_path / "use.o")], check=True)


def test_python_structs_match_c_layout(tmp_path):
    """Every ctypes mirror of a header struct has the C size and field offsets."""
    import gpuprm_amd
    pairs = [("gprm_edge_stats", gpuprm_amd.EdgeStats), ("gprm_profile_stats", gpuprm_amd.ProfileStats),
             ("gprm_prm_params", gpuprm_amd.PrmParams), ("gprm_prm_result", gpuprm_amd.PrmResult),
             ("gprm_launch_shape", gpuprm_amd.LaunchShape), ("gprm_bucket_info", gpuprm_amd.BucketInfo)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gprm.h"', 'int main(void){']
    for cname, py in pairs:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for cname, py in pairs:
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)


def _ctx(path):
    import gpuprm_amd
    L = gpuprm_amd.lib()
    c = C.c_void_p()
    rc = L.gprm_create_from_dir(path.encode(), 0.01, 64, C.byref(c))
    return L, rc, c


def test_missing_scene_is_an_error(tmp_path):
    L, rc, c = _ctx(str(tmp_path / "nope"))
    assert rc == GPRM_E_SCENE and not c.value
    assert b"ndof.bin" in L.gprm_last_error(None)


def test_corrupt_scene_is_an_error(workload_dir, tmp_path):
    import shutil
    d = str(tmp_path / "bad")
    shutil.copytree(workload_dir(1), d)
    # a vertex without edges (robotconfigspace.cpp:45-49 asserts this)
    cnt = np.fromfile(d + "/polys/poly0/cnt.bin", np.int32)
    cnt[0] = 0
    cnt.tofile(d + "/polys/poly0/cnt.bin")
    L, rc, c = _ctx(d)
    assert rc == GPRM_E_SCENE
    # truncated vertex file (the reference's read_file reads short silently)
    shutil.rmtree(d)
    shutil.copytree(workload_dir(1), d)
    open(d + "/polys/poly3/vertices.bin", "wb").write(b"\0" * 8)
    L, rc, c = _ctx(d)
    assert rc == GPRM_E_SCENE


def test_unsorted_pairs_are_rejected(workload_dir, tmp_path):
    import shutil
    d = str(tmp_path / "unsorted")
    shutil.copytree(workload_dir(1), d)
    f = np.fromfile(d + "/pairs/from.bin", np.int32)
    t = np.fromfile(d + "/pairs/to.bin", np.int32)
    f[[0, -1]] = f[[-1, 0]]
    t[[0, -1]] = t[[-1, 0]]
    f.tofile(d + "/pairs/from.bin")
    t.tofile(d + "/pairs/to.bin")
    L, rc, c = _ctx(d)
    assert rc == GPRM_E_SCENE


def test_calls_before_init_and_bad_arguments(workload_dir):
    import gpuprm_amd
    L, rc, c = _ctx(workload_dir(1))
    assert rc == 0 and c.value
    q = np.zeros(4, np.float32)
    out = C.c_int(7)
    FP = C.POINTER(C.c_float)
    assert L.gprm_indicator1(c, q.ctypes.data_as(FP), C.byref(out)) == GPRM_E_STATE
    assert L.gprm_indicator2_wait(c, 12345) == GPRM_E_STATE
    assert L.gprm_init(c, 1, 1) == GPRM_E_INVALID  # rank must be < size
    # host-side queries work without a device
    assert L.gprm_dim(c) == 4
    assert abs(L.gprm_deltaq(c) - 0.01) < 1e-9
    assert L.gprm_min(c, 0) == np.float32(-np.pi) and L.gprm_max(c, 3) == np.float32(np.pi)
    assert np.isnan(L.gprm_min(c, 9))
    res = np.zeros(3, np.int32)
    qb = np.array([0, 0, 4, 0, 0, 0, 0, 0, 0, -4, 0, 0], np.float32)  # items 1 and 0 out of bounds
    assert L.gprm_check_boundaries(c, qb.ctypes.data_as(FP), res.ctypes.data_as(C.POINTER(C.c_int)), 3, 3) == 0
    assert res.tolist() == [1, 0, 1]
    assert L.gprm_check_boundaries(c, qb.ctypes.data_as(FP), res.ctypes.data_as(C.POINTER(C.c_int)), -1, 3) \
        == GPRM_E_INVALID
    L.gprm_destroy(c)


def test_init_without_device_reports_hip_error(workload_dir):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    L, rc, c = _ctx(workload_dir(1))
    assert L.gprm_init(c, 0, 1) == GPRM_E_HIP
    L.gprm_destroy(c)


def test_scene_generate_errors(tmp_path):
    import gpuprm_amd
    L = gpuprm_amd.lib()
    assert L.gprm_scene_generate(4, 1, str(tmp_path / "x").encode()) == GPRM_E_INVALID
    assert L.gprm_scene_generate(1, 1, None) == GPRM_E_INVALID


def test_python_mirror_raises_on_errors(tmp_path):
    import gpuprm_amd
    with pytest.raises(gpuprm_amd.GprmError) as e:
        gpuprm_amd.RobotConfigspace(str(tmp_path / "missing"))
    assert e.value.code == GPRM_E_SCENE


def test_cpp_dropin_header(workload_dir, tmp_path):
    """RobotConfigspaceGPU<ndof> compiles against include/gprm_configspace.hpp,
    links libgprm.so and answers the host-side queries of Configspace<ndof>."""
    import gpuprm_amd
    src = tmp_path / "dropin.cpp"
    src.write_text(r'''
#include "gprm_configspace.hpp"
#include <cstdio>
int run(Configspace<4>* space) {  // a planner only sees the interface
  float q[4] = {0, 0, 5, 0};
  return space->dim() * 100 + space->check_boundaries(q) * 10 + (space->min(0) < 0);
}
int main(int argc, char** argv) {
  RobotConfigspaceGPU<4> cs(argv[1], 0.01f, 64);
  int r = run(&cs);
  std::printf("%d %d\n", r, cs.init(0, 1));
  try { RobotConfigspaceGPU<6> wrong(argv[1], 0.01f, 64); } catch (const GprmException& e) { std::printf("wrong ndof %d\n", e.code); }
  try { RobotConfigspaceGPU<4> missing("/nonexistent", 0.01f, 64); } catch (const GprmException& e) { std::printf("missing %d\n", e.code); }
  return 0;
}
''')
    exe = tmp_path / "dropin"
    libdir = os.path.dirname(gpuprm_amd.LIB_PATH)
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", libdir, "-lgprm", "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([str(exe), workload_dir(1)], capture_output=True, text=True, check=True).stdout.split("\n")
    r, init_rc = map(int, out[0].split())
    assert r == 411
    assert init_rc in (0, GPRM_E_HIP)
    assert out[1] == "wrong ndof %d" % GPRM_E_SCENE
    assert out[2] == "missing %d" % GPRM_E_SCENE


# The interface of the reference's Configspace<ndof> (prm/src/lib/prm/
# configspace.hpp:11-50) restated with the property that matters here: it
# declares no virtual destructor.
_BASE_WITHOUT_VIRTUAL_DTOR = r'''
template <int ndof> class Configspace {
 public:
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
'''

_USE_DROPIN = r'''
#define GPRM_HAVE_CONFIGSPACE
#include "gprm_configspace.hpp"
int use(Configspace<4>* s) { float q[4] = {0, 0, 0, 0}; return s->dim() + s->check_boundaries(q); }
template class RobotConfigspaceGPU<4>;
template class RobotConfigspaceGPU<7>;
'''


@pytest.mark.parametrize("std", ["c++11", "c++17"])
def test_cpp_dropin_against_base_without_virtual_destructor(tmp_path, std):
    """With GPRM_HAVE_CONFIGSPACE the drop-in derives from the caller's own
    Configspace<ndof>, which (like the reference's) has no virtual
    destructor; every member must still instantiate."""
    src = tmp_path / "base.cpp"
    src.write_text(_BASE_WITHOUT_VIRTUAL_DTOR + _USE_DROPIN)
    subprocess.run(["g++", "-std=" + std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)


REFERENCE_CONFIGSPACE = "/root/reference/prm/src/lib/prm/configspace.hpp"


@pytest.mark.skipif(not os.path.exists(REFERENCE_CONFIGSPACE), reason="reference sources not present")
def test_cpp_dropin_against_reference_header(tmp_path):
    """The same against the reference's own header, where it is present."""
    src = tmp_path / "ref.cpp"
    src.write_text('#include "%s"\n' % REFERENCE_CONFIGSPACE + _USE_DROPIN)
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)


def test_tuning_table_file():
    """The committed per-bucket table names only shapes the kernel has and the
    library's bucket ranges."""
    import json
    import gpuprm_amd
    t = json.load(open(gpuprm_amd.SceneBatch.TUNING_FILE))
    ranges = {(0, 512), (512, 1024), (1024, 2048), (2048, 1 << 30)}
    assert t["buckets"]
    for row in t["buckets"]:
        assert (row["min_vertices"], row["max_vertices"]) in ranges
        assert (row["block"], row["waves"]) in ((256, 5), (256, 4), (1024, 4))
        assert 1 <= row["pairs_per_unit"] <= 256
