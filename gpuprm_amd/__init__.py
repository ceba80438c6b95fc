"""gpuprm_amd -- MI355X-native exact collision checking for PRM planners.

Python mirror of GPUPRM's configuration-space interface
(prm/src/lib/prm/configspace.hpp:11-50, implemented by RobotConfigspace,
prm/src/lib/robotspace/robotconfigspace.h:26-176) over the C ABI of
``include/gprm.h`` (``libgprm.so``: HIP kernels for gfx950).

The native library is required: importing succeeds without it, but every
constructor raises ``RuntimeError`` if ``libgprm.so`` was not built, so there
is never a silent CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Optional

import numpy as np

__all__ = ["RobotConfigspace", "Surrogate", "GprmError", "EdgeStats", "ProfileStats", "PrmParams", "PrmResult", "lib",
           "generate_scene", "LIB_PATH", "LaunchShape",
           "BucketInfo", "SceneBatch", "radius_neighbours",
           "read_prm_output", "canonical_edges"]

# GPRM_LIBRARY names another build of the same sources (the sanitized host
# build of build.build_sanitized(), for CPU-side tests)
LIB_PATH = os.environ.get("GPRM_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgprm.so")

GPRM_OK = 0
_ERRORS = {
    -1: "GPRM_E_INVALID",
    -2: "GPRM_E_SCENE",
    -3: "GPRM_E_HIP",
    -4: "GPRM_E_CAPACITY",
    -5: "GPRM_E_REQUEST",
    -6: "GPRM_E_STATE",
}


class GprmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (%d): %s" % (_ERRORS.get(code, "error"), code, msg))
        self.code = code


class EdgeStats(C.Structure):
    _fields_ = [
        ("samples_total", C.c_longlong),
        ("samples_scheduled", C.c_longlong),
        ("max_testnum", C.c_int),
        ("rounds", C.c_int),
    ]


class ProfileStats(C.Structure):
    _fields_ = [
        ("sva_launches", C.c_longlong),
        ("sva_ms", C.c_double),
        ("fk_launches", C.c_longlong),
        ("fk_ms", C.c_double),
        ("pairs_separated", C.c_longlong),
        ("pairs_colliding", C.c_longlong),
        ("units_dropped", C.c_longlong),
        ("units_done", C.c_longlong),
        ("iterations", C.c_longlong),
        ("bound_separated", C.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PrmParams(C.Structure):
    _fields_ = [
        ("num", C.c_int),
        ("D", C.c_float),
        ("max_nodes", C.c_int),
        ("max_steps", C.c_int),
        ("seed", C.c_ulonglong),
        ("nbuf", C.c_int),
        ("surrogate", C.c_void_p),
        ("bias_sampling", C.c_int),
        ("order_edges", C.c_int),
        ("edge_budget", C.c_int),
        ("output_dir", C.c_char_p),
        ("workers", C.c_int),
    ]


class LaunchShape(C.Structure):
    _fields_ = [
        ("block", C.c_int),
        ("waves", C.c_int),
        ("pairs_per_unit", C.c_int),
        ("grid", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BucketInfo(C.Structure):
    _fields_ = [
        ("min_vertices", C.c_int),
        ("max_vertices", C.c_int),
        ("scenes", C.c_int),
        ("lds_bytes", C.c_int),
        ("shape", LaunchShape),
    ]


class PrmResult(C.Structure):
    _fields_ = [
        ("found", C.c_int),
        ("steps", C.c_int),
        ("path_len", C.c_int),
        ("nodes_left", C.c_longlong),
        ("nodes_right", C.c_longlong),
        ("point_checks", C.c_longlong),
        ("edges_checked", C.c_longlong),
        ("edges_free", C.c_longlong),
        ("edge_samples", C.c_longlong),
        ("model_predictions", C.c_longlong),
        ("candidates_skipped", C.c_longlong),
        ("edges_skipped", C.c_longlong),
        ("ms_total", C.c_double),
        ("ms_sampling", C.c_double),
        ("ms_near", C.c_double),
        ("ms_edges", C.c_double),
        ("ms_insert", C.c_double),
        ("ms_model", C.c_double),
        ("ms_wait", C.c_double),
        ("ms_final", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)


def lib():
    """Load libgprm.so (raises RuntimeError when it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libgprm.so not found at %s; build it with `python -m gpuprm_amd.build` "
            "(there is no CPU fallback)" % LIB_PATH
        )
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7).
    # Loading torch first makes libgprm.so bind to that already-loaded runtime,
    # so torch tensors and this library share one HIP runtime in the process;
    # two runtimes in one process cannot both open the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.gprm_create_from_dir.argtypes = [C.c_char_p, C.c_float, C.c_int, C.POINTER(vp)]
    L.gprm_create.argtypes = [C.c_int, _FP, _FP, _FP, _FP, _IP, C.c_int, _IP, _IP, _IP, _FP, _IP, _IP, _IP,
                              _IP, _IP, C.c_int, _FP, _FP, C.c_float, C.c_int, C.POINTER(vp)]
    L.gprm_init.argtypes = [vp, C.c_int, C.c_int]
    L.gprm_destroy.argtypes = [vp]
    L.gprm_destroy.restype = None
    L.gprm_indicator1.argtypes = [vp, _FP, _IP]
    L.gprm_indicator.argtypes = [vp, _FP, _IP, C.c_int, C.c_int]
    L.gprm_indicator2.argtypes = [vp, _FP, _FP, _IP, C.c_int, C.c_int]
    L.gprm_indicator2_async.argtypes = [vp, _FP, _FP, _IP, C.c_int, C.c_int, _IP]
    L.gprm_indicator2_wait.argtypes = [vp, C.c_int]
    L.gprm_indicator2_compressed.argtypes = [vp, _FP, C.c_int, _FP, _IP, _IP, _IP, C.c_int]
    L.gprm_check_boundaries.argtypes = [vp, _FP, _IP, C.c_int, C.c_int]
    L.gprm_check_boundaries1.argtypes = [vp, _FP, _IP]
    for f in ("gprm_min", "gprm_max"):
        getattr(L, f).argtypes = [vp, C.c_int]
        getattr(L, f).restype = C.c_float
    L.gprm_deltaq.argtypes = [vp]
    L.gprm_deltaq.restype = C.c_float
    L.gprm_dim.argtypes = [vp]
    L.gprm_numthreads_all.argtypes = [vp]
    L.gprm_numthreads_all.restype = C.c_longlong
    L.gprm_indicator_dev.argtypes = [vp, vp, vp, C.c_int, C.c_longlong, vp]
    L.gprm_indicator_dev_noalloc.argtypes = [vp, vp, vp, C.c_int, C.c_longlong, vp]
    L.gprm_kinematics_dev.argtypes = [vp, vp, vp, C.c_int, C.c_longlong, vp]
    L.gprm_indicator2_dev.argtypes = [vp, vp, vp, vp, C.c_int, C.c_longlong, vp, C.POINTER(EdgeStats)]
    L.gprm_last_error.argtypes = [vp]
    L.gprm_last_error.restype = C.c_char_p
    L.gprm_scene_generate.argtypes = [C.c_int, C.c_ulonglong, C.c_char_p]
    L.gprm_prm_solve.argtypes = [vp, _FP, _FP, C.POINTER(PrmParams), C.POINTER(PrmResult), _FP, C.c_int]
    L.gprm_surrogate_create.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
    L.gprm_surrogate_destroy.argtypes = [vp]
    L.gprm_surrogate_destroy.restype = None
    L.gprm_surrogate_param_count.argtypes = []
    L.gprm_surrogate_get_params.argtypes = [vp, _FP]
    L.gprm_surrogate_set_params.argtypes = [vp, _FP]
    L.gprm_surrogate_gradients.argtypes = [vp, vp, C.c_longlong, vp, C.c_int, _FP, _FP]
    L.gprm_surrogate_train_step.argtypes = [vp, vp, C.c_longlong, vp, C.c_int, C.c_float, _FP]
    L.gprm_surrogate_predict.argtypes = [vp, vp, C.c_int, C.c_longlong, vp, vp]
    L.gprm_profile_enable.argtypes = [vp, C.c_int]
    L.gprm_profile_read.argtypes = [vp, C.POINTER(ProfileStats)]
    L.gprm_get_launch_shape.argtypes = [vp, C.POINTER(LaunchShape)]
    L.gprm_set_launch_shape.argtypes = [vp, C.POINTER(LaunchShape)]
    L.gprm_autotune.argtypes = [vp, vp, vp, C.c_int, C.c_longlong, C.c_int, C.POINTER(LaunchShape), _FP]
    L.gprm_radius_neighbours.argtypes = [C.c_int, vp, C.c_int, C.c_longlong, vp, C.c_int, C.c_longlong, _FP, _FP,
                                         C.c_float, vp, vp, vp, C.c_int, C.POINTER(C.c_longlong), vp]
    PP = C.POINTER(vp)
    LL = C.POINTER(C.c_longlong)
    L.gprm_batch_create.argtypes = [PP, C.c_int, PP]
    L.gprm_batch_destroy.argtypes = [vp]
    L.gprm_batch_destroy.restype = None
    L.gprm_batch_bucket_info.argtypes = [vp, C.c_int, C.POINTER(BucketInfo)]
    L.gprm_batch_set_bucket_shape.argtypes = [vp, C.c_int, C.POINTER(LaunchShape)]
    L.gprm_batch_indicator_dev.argtypes = [vp, PP, LL, _IP, PP, C.c_int, _FP]
    L.gprm_batch_indicator2_dev.argtypes = [vp, PP, PP, LL, _IP, PP, C.c_int, _FP, C.POINTER(C.c_longlong)]
    L.gprm_batch_autotune.argtypes = [vp, PP, LL, _IP, PP, C.c_int]
    L.gprm_batch_last_error.argtypes = [vp]
    L.gprm_batch_last_error.restype = C.c_char_p
    _lib = L
    return L


def _check(rc: int, ctx=None) -> int:
    if rc < 0:
        msg = lib().gprm_last_error(ctx)
        raise GprmError(rc, msg.decode() if msg else "")
    return rc


def radius_neighbours(nodes, queries, mins, maxs, D, cap=None):
    """GPU radius search (gprm_radius_neighbours): nodes (ndof, n) and queries
    (ndof, nq) float32 device tensors, contiguous.  Returns (query, node, d2)
    numpy arrays of every pair closer than D, sorted by (query, node)."""
    import torch
    nd, n = int(nodes.shape[0]), int(nodes.shape[1])
    nq = int(queries.shape[1])
    cap = int(cap if cap is not None else max(1, n * 8))
    oq = torch.empty(cap, dtype=torch.int32, device=nodes.device)
    on = torch.empty(cap, dtype=torch.int32, device=nodes.device)
    od = torch.empty(cap, dtype=torch.float32, device=nodes.device)
    tot = C.c_longlong(0)
    rc = lib().gprm_radius_neighbours(nd, C.c_void_p(nodes.data_ptr()), n, max(1, n), C.c_void_p(queries.data_ptr()),
                                      nq, max(1, nq), _f32(mins).ctypes.data_as(_FP), _f32(maxs).ctypes.data_as(_FP),
                                      float(D), C.c_void_p(oq.data_ptr()), C.c_void_p(on.data_ptr()),
                                      C.c_void_p(od.data_ptr()), cap, C.byref(tot), None)
    _check(rc)
    k = min(cap, tot.value)
    if tot.value > cap:
        raise GprmError(-4, "%d pairs exceed cap %d" % (tot.value, cap))
    q, nn, d = oq[:k].cpu().numpy(), on[:k].cpu().numpy(), od[:k].cpu().numpy()
    o = np.lexsort((nn, q))
    return q[o], nn[o], d[o]


def canonical_edges(A, B):
    """Orient edges as the planner checks them: for rows A[k], B[k] (n, ndof)
    the lexicographically smaller configuration (at the first differing
    coordinate) becomes the start.  The samples of an edge depend on its
    direction at the last bit, so a tree or path edge is re-checked exactly as
    it was decided only in this orientation.  Returns (starts, ends)."""
    A = np.asarray(A, np.float32)
    B = np.asarray(B, np.float32)
    ne = A != B
    first = np.where(ne.any(axis=1), ne.argmax(axis=1), 0)
    k = np.arange(len(A))
    swap = B[k, first] < A[k, first]
    S = np.where(swap[:, None], B, A)
    E = np.where(swap[:, None], A, B)
    return S, E


def read_prm_output(path, ndof):
    """Read a planner output directory the way prm/matlab/lib/prmoutputread.m
    does (0-based indices here): node arrays of both trees, their edges (each
    once, from < to), the connection nodes and the path from goal to start."""
    def rd(name, dt):
        return np.fromfile(os.path.join(path, name), dtype=dt)
    out = {}
    for side in ("l", "r"):
        g = "graph" + side
        Q = rd(g + "/qstoragec.bin", np.float32).reshape(-1, ndof)
        ef, et = rd(g + "/edgesfromc.bin", np.int32), rd(g + "/edgestoc.bin", np.int32)
        keep = ef < et
        out["Q" + side] = Q
        out["E" + side] = np.stack([ef[keep], et[keep]], axis=1)
    out["conl"] = int(rd("graphl/endc.bin", np.int32)[0])
    out["conr"] = int(rd("graphr/startc.bin", np.int32)[0])
    out["connected"] = int(rd("connection.bin", np.int32)[0])
    if out["connected"]:
        out["pathl"] = rd("graphl/pathc.bin", np.int32)
        out["pathr"] = rd("graphr/pathc.bin", np.int32)
        out["Qpath"] = np.concatenate([out["Qr"][out["pathr"]], out["Ql"][out["pathl"]]], axis=0)
    return out


class SceneBatch:
    """Checks of many scenes as one launch sequence (gprm_batch_*): point
    checks (one kinematics launch, one persistent separating-vector launch per
    vertex-count bucket, one finalize launch) and edge checks (one prologue,
    then rounds of multi-scene compaction, kinematics and separating-vector
    launches); optionally replayed from a hipGraph.  `spaces` are initialised
    RobotConfigspace objects on one device with one ndof; they must outlive
    the batch."""

    BUCKETS = 4
    # per-bucket shapes measured by the autotuner on MI355X (see the file)
    TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")

    def __init__(self, spaces, tuning="default"):
        """tuning: "default" applies the committed MI355X table, None keeps
        the library's LDS-based defaults, or a table of tuning() rows."""
        self.spaces = list(spaces)
        arr = (C.c_void_p * len(self.spaces))(*[s._ctx.value for s in self.spaces])
        self._h = C.c_void_p()
        rc = lib().gprm_batch_create(arr, len(self.spaces), C.byref(self._h))
        if rc < 0:
            raise GprmError(rc, (lib().gprm_batch_last_error(None) or b"").decode())
        if tuning == "default":
            with open(self.TUNING_FILE) as f:
                tuning = json.load(f)["buckets"]
        if tuning:
            self.apply_tuning(tuning)

    def _ok(self, rc):
        if rc < 0:
            raise GprmError(rc, (lib().gprm_batch_last_error(self._h) or b"").decode())

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gprm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bucket_info(self):
        out = []
        for i in range(self.BUCKETS):
            b = BucketInfo()
            self._ok(lib().gprm_batch_bucket_info(self._h, i, C.byref(b)))
            out.append({"min_vertices": b.min_vertices, "max_vertices": b.max_vertices, "scenes": b.scenes,
                        "lds_bytes": b.lds_bytes, "shape": b.shape.as_dict()})
        return out

    def set_bucket_shape(self, bucket: int, block: int, waves: int, pairs_per_unit: int) -> None:
        sh = LaunchShape(int(block), int(waves), int(pairs_per_unit), 0)
        self._ok(lib().gprm_batch_set_bucket_shape(self._h, int(bucket), C.byref(sh)))

    def tuning(self):
        """The per-bucket shapes, as a table that apply_tuning() accepts."""
        return [{"min_vertices": b["min_vertices"], "max_vertices": b["max_vertices"],
                 "block": b["shape"]["block"], "waves": b["shape"]["waves"],
                 "pairs_per_unit": b["shape"]["pairs_per_unit"]} for b in self.bucket_info() if b["scenes"]]

    def apply_tuning(self, table) -> None:
        """Set the shape of every bucket in use whose vertex range a row of
        `table` names (rows of tuning()).  A row whose shape does not fit the
        bucket's scenes (LDS) is skipped and the bucket keeps its shape."""
        for i, b in enumerate(self.bucket_info()):
            if not b["scenes"]:
                continue
            for row in table:
                if (row["min_vertices"], row["max_vertices"]) == (b["min_vertices"], b["max_vertices"]):
                    try:
                        self.set_bucket_shape(i, row["block"], row["waves"], row["pairs_per_unit"])
                    except GprmError:
                        pass

    def _arrays(self, qs, res):
        S = len(self.spaces)
        assert len(qs) == S and len(res) == S
        qa = (C.c_void_p * S)(*[q.data_ptr() for q in qs])
        ra = (C.c_void_p * S)(*[r.data_ptr() for r in res])
        n = (C.c_int * S)(*[int(q.shape[1]) for q in qs])
        off = (C.c_longlong * S)(*[max(1, int(q.shape[1])) for q in qs])
        return qa, off, n, ra

    def indicator_dev(self, qs, res, use_graph: bool = False) -> np.ndarray:
        """qs[s]: float32 device tensor (ndof, n_s) contiguous; res[s]: int32
        device tensor (n_s,).  Returns each scene's completion time in ms
        after the separating-vector launch started (device clock)."""
        qa, off, n, ra = self._arrays(qs, res)
        done = np.zeros(len(self.spaces), np.float32)
        self._ok(lib().gprm_batch_indicator_dev(self._h, qa, off, n, ra, int(bool(use_graph)),
                                                done.ctypes.data_as(_FP)))
        return done

    def indicator2_dev(self, qs, qe, res, use_graph: bool = False):
        """Edge checks: qs[s], qe[s] float32 device tensors (ndof, n_s)
        contiguous (start and end configurations); res[s] int32 device tensor
        (n_s,), nonzero where the edge collides.  Returns (each scene's
        completion time in ms after the first separating-vector block started,
        the reference's sample count sum(testnum))."""
        S = len(self.spaces)
        assert len(qs) == S and len(qe) == S and len(res) == S
        for a, b in zip(qs, qe):
            assert a.shape == b.shape
        qa, off, n, ra = self._arrays(qs, res)
        ea = (C.c_void_p * S)(*[q.data_ptr() for q in qe])
        done = np.zeros(S, np.float32)
        total = C.c_longlong(0)
        self._ok(lib().gprm_batch_indicator2_dev(self._h, qa, ea, off, n, ra, int(bool(use_graph)),
                                                 done.ctypes.data_as(_FP), C.byref(total)))
        return done, int(total.value)

    def autotune(self, qs, res, reps: int = 2):
        qa, off, n, ra = self._arrays(qs, res)
        self._ok(lib().gprm_batch_autotune(self._h, qa, off, n, ra, int(reps)))
        return self.tuning()


def generate_scene(workload: int, seed: int, out_dir: str) -> str:
    """Write the seeded scene of a benchmark workload (1, 2, 3, 5 or 6) to out_dir."""
    _check(lib().gprm_scene_generate(int(workload), int(seed), out_dir.encode()))
    return out_dir


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Surrogate:
    """Collision-probability MLP (7 -> 256 -> 256 -> 256 -> 1) trained on the
    GPU by the library's own BF16 MFMA forward/backward kernels and Adam.  It
    ranks configurations; the exact checker decides them."""

    def __init__(self, ndof: int = 7, batch: int = 4096, seed: int = 0):
        self._h = C.c_void_p()
        _check(lib().gprm_surrogate_create(int(ndof), int(batch), int(seed), C.byref(self._h)))
        self.batch = batch
        self.ndof = ndof

    @property
    def handle(self):
        return self._h.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gprm_surrogate_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.zeros(lib().gprm_surrogate_param_count(), np.float32)
        _check(lib().gprm_surrogate_get_params(self._h, out.ctypes.data_as(_FP)))
        return out

    def set_params(self, p) -> None:
        pa = _f32(p)
        assert pa.size == lib().gprm_surrogate_param_count()
        _check(lib().gprm_surrogate_set_params(self._h, pa.ctypes.data_as(_FP)))

    def gradients(self, q, codes, offset: Optional[int] = None):
        """q: float32 device tensor (ndof, >= n), component i of row k at
        q[k + offset*i] (offset defaults to q.shape[1]); codes: int32 device
        tensor (n,).  Returns (grads, loss)."""
        n = int(codes.shape[0])
        g = np.zeros(lib().gprm_surrogate_param_count(), np.float32)
        loss = C.c_float(0)
        _check(lib().gprm_surrogate_gradients(self._h, C.c_void_p(q.data_ptr()), int(offset or q.shape[1]),
                                              C.c_void_p(codes.data_ptr()), n, g.ctypes.data_as(_FP),
                                              C.byref(loss)))
        return g, loss.value

    def train_step(self, q, codes, lr: float = 1e-3, offset: Optional[int] = None, want_loss: bool = True):
        """One Adam step on the n = len(codes) rows of q; returns the batch
        loss, or None with want_loss=False (no read-back: the step stays
        queued on the device)."""
        n = int(codes.shape[0])
        loss = C.c_float(0)
        _check(lib().gprm_surrogate_train_step(self._h, C.c_void_p(q.data_ptr()), int(offset or q.shape[1]),
                                               C.c_void_p(codes.data_ptr()), n, float(lr),
                                               C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def predict(self, q, out, stream=None, offset: Optional[int] = None) -> None:
        """Probabilities of the n = len(out) rows of q (SoA, offset defaults
        to q.shape[1]) into the float32 device tensor out."""
        n = int(out.shape[0])
        _check(lib().gprm_surrogate_predict(self._h, C.c_void_p(q.data_ptr()), n, int(offset or q.shape[1]),
                                            C.c_void_p(out.data_ptr()), C.c_void_p(stream)))


class RobotConfigspace:
    """Mirror of RobotConfigspace<ndof> (robotconfigspace.h:26-176).

    ``RobotConfigspace(path, dq, nbuf)`` loads a scene directory in the
    reference's format (robotconfigspace.cpp:100-208).  Methods keep the
    reference's names and argument meaning; they return 0 on success like the
    reference and raise ``GprmError`` where the reference would abort.
    """

    def __init__(self, path: Optional[str] = None, dq: float = 0.01, nbuf: int = 2048, *, arrays=None):
        L = lib()
        self._ctx = C.c_void_p()
        if path is not None:
            _check(L.gprm_create_from_dir(path.encode(), float(dq), int(nbuf), C.byref(self._ctx)))
        elif arrays is not None:
            self._create_from_arrays(arrays, dq, nbuf)
        else:
            raise ValueError("need a scene path or arrays")
        self.ndof = L.gprm_dim(self._ctx)
        self._reqs = {}

    def _create_from_arrays(self, s, dq, nbuf):
        # s: dict with the flattened constructor arguments of gprm_create
        keep = {k: (_i32(v) if k in ("types", "sys", "n", "m", "dsp", "cnt", "dest", "from", "to") else _f32(v))
                for k, v in s.items() if k not in ("ndof",)}
        self._keep = keep
        fp = lambda k: keep[k].ctypes.data_as(_FP)
        ip = lambda k: keep[k].ctypes.data_as(_IP)
        _check(lib().gprm_create(
            int(s["ndof"]), fp("a"), fp("alpha"), fp("q"), fp("d"), ip("types"), len(keep["sys"]), ip("sys"),
            ip("n"), ip("m"), fp("vertices"), ip("dsp"), ip("cnt"), ip("dest"), ip("from"), ip("to"),
            len(keep["from"]), fp("mins"), fp("maxs"), float(dq), int(nbuf), C.byref(self._ctx)))

    # -- lifecycle ---------------------------------------------------------
    def init(self, ressource_rank: int = 0, ressource_size: int = 1) -> int:
        """Configspace::init (configspace.hpp:16)."""
        return _check(lib().gprm_init(self._ctx, int(ressource_rank), int(ressource_size)), self._ctx)

    def close(self) -> None:
        if getattr(self, "_ctx", None) and self._ctx.value:
            lib().gprm_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- indicator functions ---------------------------------------------
    def indicator(self, q, res=None, N: Optional[int] = None, offset: Optional[int] = None):
        """indicator(q) -> code (configspace.hpp:22), or the batched form
        indicator(q, res, N, offset) -> 0 (configspace.hpp:17-20), SoA with
        component i of item k at q[k + offset*i]."""
        L = lib()
        qa = _f32(q)
        if res is None:
            out = C.c_int(0)
            _check(L.gprm_indicator1(self._ctx, qa.ctypes.data_as(_FP), C.byref(out)), self._ctx)
            return out.value
        if N is None or offset is None:
            raise ValueError("batched indicator needs N and offset")
        self._check_soa(qa, N, offset)
        r = self._out_i32(res, N)
        _check(L.gprm_indicator(self._ctx, qa.ctypes.data_as(_FP), r.ctypes.data_as(_IP), int(N), int(offset)),
               self._ctx)
        if r is not res:
            res[:N] = r[:N]
        return 0

    def indicator_batch(self, Q) -> np.ndarray:
        """Convenience: Q of shape (ndof, N) -> codes (N,)."""
        Qa = _f32(Q)
        N = Qa.shape[1]
        res = np.zeros(N, np.int32)
        self.indicator(Qa.ravel(), res, N, N)
        return res

    def indicator2(self, qs, qe, res, N: int, offset: int) -> int:
        """indicator2 (configspace.hpp:23-26)."""
        req = self.indicator2_async(qs, qe, res, N, offset)
        return self.indicator2_async_wait(req)

    def indicator2_async(self, qs, qe, res, N: int, offset: int) -> int:
        """indicator2_async (configspace.hpp:27); returns the request id."""
        qsa, qea = _f32(qs), _f32(qe)
        self._check_soa(qsa, N, offset)
        self._check_soa(qea, N, offset)
        r = self._out_i32(res, N)
        req = C.c_int(0)
        _check(lib().gprm_indicator2_async(self._ctx, qsa.ctypes.data_as(_FP), qea.ctypes.data_as(_FP),
                                           r.ctypes.data_as(_IP), int(N), int(offset), C.byref(req)), self._ctx)
        self._reqs[req.value] = (qsa, qea, r, res)
        return req.value

    def indicator2_async_wait(self, request: int) -> int:
        """indicator2_async_wait (configspace.hpp:28)."""
        _check(lib().gprm_indicator2_wait(self._ctx, int(request)), self._ctx)
        qsa, qea, r, res = self._reqs.pop(request)
        if r is not res:
            n = min(len(res), len(r))
            res[:n] = r[:n]
        return 0

    def indicator2_compressed(self, qs, M: int, qe, res, posqe, numqe, offset: int) -> int:
        """Compressed indicator2 (configspace.hpp:30-32)."""
        qsa, qea = _f32(qs), _f32(qe)
        pa, na = _i32(posqe), _i32(numqe)
        total = int(max((pa[i] + na[i] for i in range(M)), default=0))
        r = self._out_i32(res, total)
        _check(lib().gprm_indicator2_compressed(self._ctx, qsa.ctypes.data_as(_FP), int(M), qea.ctypes.data_as(_FP),
                                                r.ctypes.data_as(_IP), pa.ctypes.data_as(_IP), na.ctypes.data_as(_IP),
                                                int(offset)), self._ctx)
        if r is not res:
            res[:total] = r[:total]
        return 0

    def check_boundaries(self, q, res=None, N: Optional[int] = None, offset: Optional[int] = None):
        """check_boundaries (configspace.hpp:33-37)."""
        qa = _f32(q)
        if res is None:
            out = C.c_int(0)
            _check(lib().gprm_check_boundaries1(self._ctx, qa.ctypes.data_as(_FP), C.byref(out)), self._ctx)
            return out.value
        self._check_soa(qa, N, offset)
        r = self._out_i32(res, N)
        _check(lib().gprm_check_boundaries(self._ctx, qa.ctypes.data_as(_FP), r.ctypes.data_as(_IP), int(N),
                                           int(offset)), self._ctx)
        if r is not res:
            res[:N] = r[:N]
        return 0

    # -- getters (configspace.hpp:39-47) -----------------------------------
    def min(self, i: int) -> float:
        return lib().gprm_min(self._ctx, int(i))

    def max(self, i: int) -> float:
        return lib().gprm_max(self._ctx, int(i))

    def deltaq(self) -> float:
        return lib().gprm_deltaq(self._ctx)

    def dim(self) -> int:
        return self.ndof

    def get_numthreads_all(self) -> int:
        return lib().gprm_numthreads_all(self._ctx)

    # -- device-resident variants (torch tensors on the GPU) -------------
    def indicator_dev(self, q, res, stream=None) -> None:
        """q: float32 device tensor (ndof, N) contiguous; res: int32 device tensor (N,)."""
        N = int(q.shape[1])
        _check(lib().gprm_indicator_dev(self._ctx, C.c_void_p(q.data_ptr()), C.c_void_p(res.data_ptr()), N,
                                        N, C.c_void_p(stream)), self._ctx)

    def kinematics_dev(self, q, frames, stream=None) -> None:
        """Kinematics<ndof>::calculate (kinematics.h:28-37) for every column of
        q (float32 device tensor (ndof, N), contiguous): frames float32 device
        tensor (N, ndof + 1, 4, 3), frame s, column c, row x/y/z."""
        N = int(q.shape[1])
        assert tuple(frames.shape) == (N, self.ndof + 1, 4, 3) and frames.is_contiguous()
        _check(lib().gprm_kinematics_dev(self._ctx, C.c_void_p(q.data_ptr()), C.c_void_p(frames.data_ptr()), N,
                                         N, C.c_void_p(stream)), self._ctx)

    def indicator2_dev(self, qs, qe, res, stream=None) -> EdgeStats:
        """qs, qe: float32 device tensors (ndof, N); res: int32 device tensor (N,)."""
        N = int(qs.shape[1])
        st = EdgeStats()
        _check(lib().gprm_indicator2_dev(self._ctx, C.c_void_p(qs.data_ptr()), C.c_void_p(qe.data_ptr()),
                                         C.c_void_p(res.data_ptr()), N, N, C.c_void_p(stream), C.byref(st)),
               self._ctx)
        return st

    # -- single-process PRM ----------------------------------------------
    def prm_solve(self, qstart, qgoal, num=2048, D=0.5, max_nodes=1 << 18, max_steps=200, seed=1,
                  nbuf=1 << 20, surrogate=None, bias_sampling=False, order_edges=False, edge_budget=0,
                  path_cap=1 << 16, output_dir=None, workers=1):
        """Bidirectional PRM from qstart to qgoal (the reference's PRMSolver
        process, prm/src/lib/prm/prmsolver.cpp:1765-2187, on one GPU).
        edge_budget=0 checks every candidate edge, as the reference does; k > 0
        checks at most k per (new node, tree), shortest first (with
        order_edges, the edges the surrogate predicts blocked last).
        output_dir: write both trees in the reference's output layout (see
        read_prm_output).  workers: steps in flight (1, or 2 as the
        reference's process 5).  Returns (PrmResult, path (n, ndof) or None)."""
        p = PrmParams(int(num), float(D), int(max_nodes), int(max_steps), int(seed), int(nbuf),
                      C.c_void_p(surrogate.handle if surrogate is not None else None),
                      int(bool(bias_sampling)), int(bool(order_edges)), int(edge_budget),
                      output_dir.encode() if output_dir else None, int(workers))
        r = PrmResult()
        path = np.zeros((path_cap, self.ndof), np.float32)
        _check(lib().gprm_prm_solve(self._ctx, _f32(qstart).ctypes.data_as(_FP), _f32(qgoal).ctypes.data_as(_FP),
                                    C.byref(p), C.byref(r), path.ctypes.data_as(_FP), int(path_cap)), self._ctx)
        return r, (path[:min(r.path_len, path_cap)].copy() if r.found else None)

    # -- kernel timing (HIP events) and work counters ---------------------
    def profile_enable(self, enable: bool = True) -> None:
        _check(lib().gprm_profile_enable(self._ctx, int(bool(enable))), self._ctx)

    def profile_read(self) -> ProfileStats:
        st = ProfileStats()
        _check(lib().gprm_profile_read(self._ctx, C.byref(st)), self._ctx)
        return st

    # -- launch shape and autotuner (no reference counterpart) -----------
    def get_launch_shape(self) -> LaunchShape:
        sh = LaunchShape()
        _check(lib().gprm_get_launch_shape(self._ctx, C.byref(sh)), self._ctx)
        return sh

    def set_launch_shape(self, block: int, waves: int, pairs_per_unit: int) -> LaunchShape:
        sh = LaunchShape(int(block), int(waves), int(pairs_per_unit), 0)
        _check(lib().gprm_set_launch_shape(self._ctx, C.byref(sh)), self._ctx)
        return self.get_launch_shape()

    def autotune(self, q, qe=None, reps: int = 3):
        """Time every kernel variant x pairs-per-unit on device tensors q (and
        qe for edges), (ndof, N) contiguous; keeps and returns the fastest
        shape and its time in ms."""
        N = int(q.shape[1])
        sh = LaunchShape()
        ms = C.c_float(0)
        _check(lib().gprm_autotune(self._ctx, C.c_void_p(q.data_ptr()),
                                   C.c_void_p(qe.data_ptr() if qe is not None else None), N, N, int(reps),
                                   C.byref(sh), C.byref(ms)), self._ctx)
        return sh, ms.value

    # -- helpers -----------------------------------------------------------
    def _check_soa(self, a, N, offset):
        if N is None or offset is None or N < 0 or offset < 0:
            raise ValueError("bad N/offset")
        if N > 0 and (self.ndof - 1) * offset + N > a.size:
            raise ValueError("array too small for N=%d offset=%d ndof=%d" % (N, offset, self.ndof))

    @staticmethod
    def _out_i32(res, N):
        if isinstance(res, np.ndarray) and res.dtype == np.int32 and res.flags.c_contiguous and res.size >= N:
            return res
        return np.zeros(max(N, 1), np.int32)
