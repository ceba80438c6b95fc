"""gprm_indicator_dev_noalloc (point checks into the library's fixed scratch
array, no allocation) through the C ABI: the same codes as
gprm_indicator_dev on two calls in a row, the second smaller, and a clear
error beyond its maximum size."""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu

GPRM_E_CAPACITY = -4


def _run(L, ctx, fn, q, n, torch):
    res = torch.full((n,), 12345, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = fn(ctx, C.c_void_p(q.data_ptr()), C.c_void_p(res.data_ptr()), n, q.shape[1], None)
    torch.cuda.synchronize()
    assert rc == 0, L.gprm_last_error(ctx)
    return res.cpu().numpy()


@pytest.mark.parametrize("wl,sizes", [(1, (65536, 1000)), (3, (40000, 777)), (2, (4096, 64))])
def test_noalloc_matches_indicator_dev(workload_dir, wl, sizes):
    import torch

    import gpuprm_amd
    d = workload_dir(wl)
    space = gpuprm_amd.RobotConfigspace(d, 0.01, 1024)
    space.init(0, 1)
    L = gpuprm_amd.lib()
    ctx = space._ctx
    o = OracleScene(d)
    rng = np.random.default_rng(77 + wl)
    for n in sizes:  # two calls in a row, the second smaller
        print("noalloc: workload %d, %d configurations" % (wl, n), flush=True)
        Q = rng.uniform(-3.3, 3.3, size=(o.ndof, n)).astype(np.float32)  # a few out of bounds
        q = torch.from_numpy(Q).cuda()
        want = _run(L, ctx, L.gprm_indicator_dev, q, n, torch)
        got = _run(L, ctx, L.gprm_indicator_dev_noalloc, q, n, torch)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (wl, n, bad[:8], got[bad[:8]], want[bad[:8]])
        sub = rng.choice(n, min(n, 2000), replace=False)
        assert np.array_equal(got[sub], o.indicator_batch(Q[:, sub].copy(), nthreads=8))


def test_noalloc_rejects_beyond_maximum(workload_dir):
    import torch

    import gpuprm_amd
    d = workload_dir(1)
    space = gpuprm_amd.RobotConfigspace(d, 0.01, 1024)
    space.init(0, 1)
    L = gpuprm_amd.lib()
    n = 65537
    q = torch.zeros((4, n), dtype=torch.float32, device="cuda")
    r = torch.zeros(n, dtype=torch.int32, device="cuda")
    rc = L.gprm_indicator_dev_noalloc(space._ctx, C.c_void_p(q.data_ptr()), C.c_void_p(r.data_ptr()), n, n, None)
    assert rc == GPRM_E_CAPACITY
    assert b"maximum" in L.gprm_last_error(space._ctx)
