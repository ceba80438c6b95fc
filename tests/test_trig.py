"""The restated glibc sinf/cosf of gpuprm_amd/csrc/glibc_trig.h, compiled for
the host, against the host's glibc sinf/cosf -- the functions the
reference's host path calls for sin/cos of a float (prm/src/lib/collision4/
cuda_head.h:25-28, geo4.h:184-187).  Exhaustive over every finite float
with |x| < 256 (both signs), plus every 97th float pattern beyond, which
exercises the large-argument reduction."""
import os
import subprocess
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = textwrap.dedent(r'''
    #include "glibc_trig.h"
    #include <math.h>
    #include <pthread.h>
    #include <stdio.h>
    #include <stdlib.h>
    #include <string.h>
    typedef struct { unsigned long long lo, hi, step, bad; } job;
    static unsigned bits(float f) { unsigned u; memcpy(&u, &f, 4); return u; }
    static void* run(void* a) {
      job* j = (job*)a;
      for (unsigned long long u = j->lo; u < j->hi; u += j->step) {
        for (int sg = 0; sg < 2; ++sg) {
          unsigned b = (unsigned)u | (sg ? 0x80000000u : 0u);
          float x; memcpy(&x, &b, 4);
          float s = sinf(x), c = cosf(x);
          float s2 = gprm_trig::sinf_ref(x), c2 = gprm_trig::cosf_ref(x);
          if (bits(s) != bits(s2) && !(s != s && s2 != s2)) { if (j->bad++ < 3) printf("sin %a %a %a\n", x, s, s2); }
          if (bits(c) != bits(c2) && !(c != c && c2 != c2)) { if (j->bad++ < 3) printf("cos %a %a %a\n", x, c, c2); }
        }
      }
      return 0;
    }
    int main(int argc, char** argv) {
      unsigned long long lo = strtoull(argv[1], 0, 0), hi = strtoull(argv[2], 0, 0), step = strtoull(argv[3], 0, 0);
      int T = 8; pthread_t th[8]; job js[8];
      for (int t = 0; t < T; ++t) {
        unsigned long long n = (hi - lo) / step;
        js[t].lo = lo + (n * t / T) * step; js[t].hi = lo + (n * (t + 1) / T) * step; js[t].step = step; js[t].bad = 0;
        pthread_create(&th[t], 0, run, &js[t]);
      }
      unsigned long long bad = 0;
      for (int t = 0; t < T; ++t) { pthread_join(th[t], 0); bad += js[t].bad; }
      printf("mismatches %llu\n", bad);
      return bad != 0;
    }
''')


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("trig")
    src = d / "trig_check.cpp"
    src.write_text(SRC)
    exe = d / "trig_check"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fno-builtin", "-I",
                    os.path.join(ROOT, "gpuprm_amd", "csrc"), str(src), "-o", str(exe), "-lm", "-lpthread"],
                   check=True)
    return str(exe)


def test_exhaustive_below_256(checker):
    # every float with |x| < 256: all joint angles, both reduction paths
    out = subprocess.run([checker, "0", str(0x43800000), "1"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout


def test_sampled_large_arguments(checker):
    # beyond 256 up to infinity: every 97th bit pattern (large-argument reduction)
    out = subprocess.run([checker, str(0x43800000), str(0x7f800000), "97"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout


def test_special_values(checker):
    # +-0, +-inf, NaN: the last three give NaN in both (compared as "both NaN")
    for b in (0x0, 0x7f800000, 0x7fc00000):
        out = subprocess.run([checker, str(b), str(b + 1), "1"], capture_output=True, text=True)
        assert out.returncode == 0, out.stdout
