"""GPU forward kinematics (gprm_kinematics_dev: the kinematics kernel of the
point checks, whose frames the separating-vector kernel reads, unpacked)
against the reference's Kinematics<ndof>::calculate (kinematics.h:28-37):
bit for bit against the frames the reference wrote into the golden fixtures,
and against the oracle's restatement on thousands of configurations,
including ones outside the joint limits and a strided (offset > N) input."""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu

GOLDEN_NAMES = ("config1", "wl1", "wl2", "wl3")


def _space(d):
    import gpuprm_amd
    s = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 16)
    s.init(0, 1)
    return s


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_frames_match_reference_golden(golden_dir, name):
    import torch
    z, d = golden_dir(name)
    s = _space(d)
    nd = s.dim()
    k = z["fk"].shape[0]
    print("kinematics: golden %s, %d configurations" % (name, k), flush=True)
    q = torch.from_numpy(np.ascontiguousarray(z["Q"][:, :k])).cuda()
    F = torch.zeros((k, nd + 1, 4, 3), dtype=torch.float32, device="cuda")
    s.kinematics_dev(q, F)
    assert np.array_equal(F.cpu().numpy().view(np.uint32), z["fk"].view(np.uint32))


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_frames_match_oracle_strided(golden_dir, name):
    import torch

    import gpuprm_amd
    z, d = golden_dir(name)
    s = _space(d)
    o = OracleScene(d)
    nd = s.dim()
    N, pad = 3001, 77  # ragged: not a multiple of 64; rows of N + pad floats
    rng = np.random.default_rng(5)
    Q = rng.uniform(-2 * np.pi, 2 * np.pi, size=(nd, N + pad)).astype(np.float32)
    print("kinematics: %s, %d strided configurations" % (name, N), flush=True)
    q = torch.from_numpy(Q).cuda()
    F = torch.zeros((N, nd + 1, 4, 3), dtype=torch.float32, device="cuda")
    rc = gpuprm_amd.lib().gprm_kinematics_dev(s._ctx, C.c_void_p(q.data_ptr()), C.c_void_p(F.data_ptr()), N,
                                              N + pad, None)
    assert rc == 0
    got = F.cpu().numpy()
    for i in range(N):
        want = o.fk(Q[:, i])[:, :, :3]
        assert np.array_equal(got[i].view(np.uint32), want.view(np.uint32)), i
