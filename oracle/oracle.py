"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
CPU path (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (gpuprm_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GPRM_ORACLE_SUFFIX selects another build of the same sources, e.g. "_san"
# for the AddressSanitizer/UndefinedBehaviorSanitizer build (make san)
_SFX = os.environ.get("GPRM_ORACLE_SUFFIX", "")
ORACLE_SO = os.path.join(HERE, "liboracle%s.so" % _SFX)
ORACLE_COUNT_SO = os.path.join(HERE, "liboracle_count%s.so" % _SFX)
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)
_VP = C.c_void_p


def _fp(a):
    return a.ctypes.data_as(_FP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def _load_oracle(path):
    if not os.path.exists(path):
        raise RuntimeError("%s missing; run `make -C oracle`" % path)
    L = C.CDLL(path)
    L.oracle_scene_load_dir.restype = _VP
    L.oracle_scene_load_dir.argtypes = [C.c_char_p, C.c_float]
    L.oracle_scene_free.argtypes = [_VP]
    L.oracle_scene_free.restype = None
    L.oracle_fk.argtypes = [_VP, _FP, _FP]
    L.oracle_fk.restype = None
    L.oracle_sva.argtypes = [_VP, C.c_int, C.c_int, _FP, _FP, _IP]
    L.oracle_check_config.argtypes = [_VP, _FP, _IP]
    L.oracle_indicator1.argtypes = [_VP, _FP]
    L.oracle_indicator_batch.argtypes = [_VP, _FP, _IP, C.c_int, C.c_int, C.c_int]
    L.oracle_check_boundaries.argtypes = [_VP, _FP, _IP, C.c_int, C.c_int]
    L.oracle_check_boundaries1.argtypes = [_VP, _FP]
    L.oracle_edge_testnum.argtypes = [_VP, _FP, _FP, C.c_int, C.c_int, _IP, C.POINTER(C.c_longlong)]
    L.oracle_edge_testnum.restype = C.c_longlong
    L.oracle_indicator2.argtypes = [_VP, _FP, _FP, _IP, C.c_int, C.c_int, C.c_int]
    L.oracle_indicator2.restype = C.c_longlong
    L.oracle_edge_sample.argtypes = [_VP, _FP, _FP, C.c_int, C.c_int, C.c_int, C.c_int, _FP]
    L.oracle_edge_sample.restype = None
    L.oracle_counters.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
    L.oracle_counters.restype = None
    L.oracle_pair_profile.argtypes = [_VP, _FP, _IP, C.POINTER(C.c_ulonglong)]
    L.oracle_counters_reset.argtypes = []
    L.oracle_counters_reset.restype = None
    return L


_libs = {}


def oracle_lib(counting: bool = False):
    key = "count" if counting else "plain"
    if key not in _libs:
        _libs[key] = _load_oracle(ORACLE_COUNT_SO if counting else ORACLE_SO)
    return _libs[key]


class OracleScene:
    """A scene loaded by the oracle from the reference's directory format."""

    # offsets into oracle_scene (oracle.h) are not needed: accessors go through C

    def __init__(self, path: str, dq: float = 0.01, counting: bool = False):
        self.L = oracle_lib(counting)
        self.ptr = self.L.oracle_scene_load_dir(path.encode(), float(dq))
        if not self.ptr:
            raise RuntimeError("oracle could not load scene %s" % path)
        self.path = path
        self.ndof = int(np.fromfile(os.path.join(path, "ndof.bin"), np.int32)[0])
        self.N = int(np.fromfile(os.path.join(path, "polys/N.bin"), np.int32)[0])
        self.sys = np.fromfile(os.path.join(path, "polys/sys.bin"), np.int32)
        self.pairs_from = np.fromfile(os.path.join(path, "pairs/from.bin"), np.int32)
        self.pairs_to = np.fromfile(os.path.join(path, "pairs/to.bin"), np.int32)

    def __del__(self):
        try:
            if self.ptr:
                self.L.oracle_scene_free(self.ptr)
        except Exception:
            pass

    def fk(self, q) -> np.ndarray:
        """(ndof+1, 4 columns, 4) frames; only x,y,z of each column are meaningful."""
        qa = np.ascontiguousarray(q, np.float32)
        out = np.zeros((self.ndof + 1) * 16, np.float32)
        self.L.oracle_fk(self.ptr, _fp(qa), _fp(out))
        return out.reshape(self.ndof + 1, 4, 4)

    def sva(self, q, k0: int, k1: int):
        T = self.fk(q)
        it = C.c_int(0)
        tp = np.ascontiguousarray(T[self.sys[k0]])
        tq = np.ascontiguousarray(T[self.sys[k1]])
        code = self.L.oracle_sva(self.ptr, k0, k1, _fp(tp), _fp(tq), C.byref(it))
        return code, it.value

    def indicator1(self, q) -> int:
        qa = np.ascontiguousarray(q, np.float32)
        return self.L.oracle_indicator1(self.ptr, _fp(qa))

    def check_config(self, q):
        qa = np.ascontiguousarray(q, np.float32)
        n = C.c_int(0)
        code = self.L.oracle_check_config(self.ptr, _fp(qa), C.byref(n))
        return code, n.value

    def indicator_batch(self, Q, nthreads: int = 1) -> np.ndarray:
        """Q (ndof, N) -> codes, indicator1 semantics per item."""
        Qa = np.ascontiguousarray(Q, np.float32)
        N = Qa.shape[1]
        res = np.zeros(max(N, 1), np.int32)
        self.L.oracle_indicator_batch(self.ptr, _fp(Qa), _ip(res), N, N, nthreads)
        return res[:N]

    def edge_testnum(self, QS, QE):
        QSa = np.ascontiguousarray(QS, np.float32)
        QEa = np.ascontiguousarray(QE, np.float32)
        N = QSa.shape[1]
        tn = np.zeros(max(N, 1), np.int32)
        tot = self.L.oracle_edge_testnum(self.ptr, _fp(QSa), _fp(QEa), N, N, _ip(tn), None)
        return tn[:N], tot

    def edge_sample(self, QS, QE, k: int, j: int, testnum: int) -> np.ndarray:
        QSa = np.ascontiguousarray(QS, np.float32)
        QEa = np.ascontiguousarray(QE, np.float32)
        out = np.zeros(self.ndof, np.float32)
        self.L.oracle_edge_sample(self.ptr, _fp(QSa), _fp(QEa), k, QSa.shape[1], j, testnum, _fp(out))
        return out

    def indicator2(self, QS, QE, nthreads: int = 1):
        """QS, QE (ndof, N) -> (codes (N,), numthreads) with the reference's host-loop semantics."""
        QSa = np.ascontiguousarray(QS, np.float32)
        QEa = np.ascontiguousarray(QE, np.float32)
        N = QSa.shape[1]
        res = np.zeros(max(N, 1), np.int32)
        tot = self.L.oracle_indicator2(self.ptr, _fp(QSa), _fp(QEa), _ip(res), N, N, nthreads)
        return res[:N], tot

    def pair_profile(self, q):
        """(codes, flops) of the pair tests the reference's loop runs for q."""
        qa = np.ascontiguousarray(q, np.float32)
        M = len(self.pairs_from)
        codes = np.zeros(M, np.int32)
        flops = np.zeros(M, np.uint64)
        n = self.L.oracle_pair_profile(self.ptr, _fp(qa), _ip(codes),
                                       flops.ctypes.data_as(C.POINTER(C.c_ulonglong)))
        return codes[:n], flops[:n]

    def counters(self):
        f, t = C.c_ulonglong(0), C.c_ulonglong(0)
        self.L.oracle_counters(C.byref(f), C.byref(t))
        return f.value, t.value

    def reset_counters(self):
        self.L.oracle_counters_reset()


_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            raise RuntimeError("reference build %s missing" % REF_SO)
        L = C.CDLL(REF_SO)
        L.ref_create_dir.restype = _VP
        L.ref_create_dir.argtypes = [C.c_char_p, C.c_float, C.c_int]
        L.ref_destroy.argtypes = [_VP]
        L.ref_destroy.restype = None
        L.ref_ndof.argtypes = [_VP]
        L.ref_indicator1.argtypes = [_VP, _FP]
        L.ref_indicator2.argtypes = [_VP, _FP, _FP, _IP, C.c_int, C.c_int]
        L.ref_fk.argtypes = [_VP, _FP, _FP]
        L.ref_fk.restype = None
        L.ref_sva.argtypes = [_VP, _FP, C.c_int, C.c_int, _IP]
        L.ref_numthreads_all.argtypes = [_VP]
        L.ref_numthreads_all.restype = C.c_longlong
        L.ref_indicator_mt.argtypes = [C.POINTER(_VP), C.c_int, _FP, _IP, C.c_int, C.c_int]
        L.ref_indicator2_mt.argtypes = [C.POINTER(_VP), C.c_int, _FP, _FP, _IP, C.c_int, C.c_int]
        _ref = L
    return _ref


class RefSpace:
    """The reference's RobotConfigspace (host path) loaded from a scene dir."""

    def __init__(self, path: str, dq: float = 0.01, nbuf: int = 1 << 20):
        self.L = ref_lib()
        self.h = self.L.ref_create_dir(path.encode(), float(dq), int(nbuf))
        if not self.h:
            raise RuntimeError("reference could not load %s" % path)
        self.ndof = self.L.ref_ndof(self.h)

    def __del__(self):
        try:
            if self.h:
                self.L.ref_destroy(self.h)
        except Exception:
            pass

    def indicator1(self, q) -> int:
        qa = np.ascontiguousarray(q, np.float32)
        return self.L.ref_indicator1(self.h, _fp(qa))

    def fk(self, q) -> np.ndarray:
        qa = np.ascontiguousarray(q, np.float32)
        out = np.zeros((self.ndof + 1) * 12, np.float32)
        self.L.ref_fk(self.h, _fp(qa), _fp(out))
        return out.reshape(self.ndof + 1, 4, 3)

    def sva(self, q, k0: int, k1: int):
        qa = np.ascontiguousarray(q, np.float32)
        it = C.c_int(0)
        code = self.L.ref_sva(self.h, _fp(qa), k0, k1, C.byref(it))
        return code, it.value

    def indicator2(self, QS, QE):
        QSa = np.ascontiguousarray(QS, np.float32)
        QEa = np.ascontiguousarray(QE, np.float32)
        N = QSa.shape[1]
        res = np.zeros(max(N, 1), np.int32)
        self.L.ref_indicator2(self.h, _fp(QSa), _fp(QEa), _ip(res), N, N)
        return res[:N]

    def numthreads_all(self) -> int:
        return self.L.ref_numthreads_all(self.h)


class RefPool:
    """nthreads reference objects for multi-threaded batch calls (the
    reference object is not thread-safe: its kinematics frames are a member)."""

    def __init__(self, path: str, nthreads: int, dq: float = 0.01):
        self.spaces = [RefSpace(path, dq) for _ in range(nthreads)]
        self.n = nthreads
        self.handles = (_VP * nthreads)(*[s.h for s in self.spaces])
        self.L = ref_lib()

    def indicator_batch(self, Q) -> np.ndarray:
        Qa = np.ascontiguousarray(Q, np.float32)
        N = Qa.shape[1]
        res = np.zeros(max(N, 1), np.int32)
        self.L.ref_indicator_mt(self.handles, self.n, _fp(Qa), _ip(res), N, N)
        return res[:N]

    def indicator2(self, QS, QE) -> np.ndarray:
        QSa = np.ascontiguousarray(QS, np.float32)
        QEa = np.ascontiguousarray(QE, np.float32)
        N = QSa.shape[1]
        res = np.zeros(max(N, 1), np.int32)
        self.L.ref_indicator2_mt(self.handles, self.n, _fp(QSa), _fp(QEa), _ip(res), N, N)
        return res[:N]
