"""Scene directory format in numpy (reference: prm/src/lib/robotspace/
robotconfigspace.cpp:100-208, prm/matlab/configwrite.m), used by the tests to
pack scenes into single .npz fixtures and to unpack them into directories.
Independent of the C++ loader/writer, so round trips check both."""
from __future__ import annotations

import os

import numpy as np

ROBOT_KEYS = ("a", "alpha", "q", "d")


def read_dir(path: str) -> dict:
    r = lambda p, t: np.fromfile(os.path.join(path, p), dtype=t)
    s = {"ndof": np.int32(r("ndof.bin", np.int32)[0])}
    for k in ROBOT_KEYS:
        s[k] = r("dh/%s.bin" % k, np.float32)
    s["types"] = r("dh/types.bin", np.int32)
    s["mins"] = r("mins.bin", np.float32)
    s["maxs"] = r("maxs.bin", np.float32)
    N = int(r("polys/N.bin", np.int32)[0])
    s["sys"] = r("polys/sys.bin", np.int32)
    s["from"] = r("pairs/from.bin", np.int32)
    s["to"] = r("pairs/to.bin", np.int32)
    n, m, v, dsp, cnt, dest = [], [], [], [], [], []
    for i in range(N):
        pd = "polys/poly%d/" % i
        sz = r(pd + "size.bin", np.int32)
        n.append(sz[0])
        m.append(sz[1])
        v.append(r(pd + "vertices.bin", np.float32))
        dsp.append(r(pd + "dsp.bin", np.int32))
        cnt.append(r(pd + "cnt.bin", np.int32))
        dest.append(r(pd + "dest.bin", np.int32))
    s["n"] = np.array(n, np.int32)
    s["m"] = np.array(m, np.int32)
    s["vertices"] = np.concatenate(v).astype(np.float32)
    s["dsp"] = np.concatenate(dsp).astype(np.int32)
    s["cnt"] = np.concatenate(cnt).astype(np.int32)
    s["dest"] = np.concatenate(dest).astype(np.int32)
    return s


def write_dir(s: dict, path: str) -> str:
    for sub in ("dh", "polys", "pairs"):
        os.makedirs(os.path.join(path, sub), exist_ok=True)
    w = lambda p, a, t: np.ascontiguousarray(a, dtype=t).tofile(os.path.join(path, p))
    w("ndof.bin", [int(s["ndof"])], np.int32)
    for k in ROBOT_KEYS:
        w("dh/%s.bin" % k, s[k], np.float32)
    w("dh/types.bin", s["types"], np.int32)
    w("mins.bin", s["mins"], np.float32)
    w("maxs.bin", s["maxs"], np.float32)
    N = len(s["n"])
    w("polys/N.bin", [N], np.int32)
    w("polys/sys.bin", s["sys"], np.int32)
    w("pairs/M.bin", [len(s["from"])], np.int32)
    w("pairs/from.bin", s["from"], np.int32)
    w("pairs/to.bin", s["to"], np.int32)
    vo = eo = 0
    for i in range(N):
        pd = "polys/poly%d" % i
        os.makedirs(os.path.join(path, pd), exist_ok=True)
        n, m = int(s["n"][i]), int(s["m"][i])
        w(pd + "/size.bin", [n, m], np.int32)
        w(pd + "/vertices.bin", s["vertices"][3 * vo:3 * (vo + n)], np.float32)
        w(pd + "/dsp.bin", s["dsp"][vo:vo + n], np.int32)
        w(pd + "/cnt.bin", s["cnt"][vo:vo + n], np.int32)
        w(pd + "/dest.bin", s["dest"][eo:eo + m], np.int32)
        vo += n
        eo += m
    return path


SCENE_KEYS = ("ndof", "a", "alpha", "q", "d", "types", "mins", "maxs", "sys", "from", "to", "n", "m",
              "vertices", "dsp", "cnt", "dest")


def scene_from_npz(z) -> dict:
    return {k: z["scene_" + k] for k in SCENE_KEYS}


def scene_to_npz_dict(s: dict) -> dict:
    return {"scene_" + k: np.asarray(s[k]) for k in SCENE_KEYS}
