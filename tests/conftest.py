"""Shared fixtures.  Tests marked `gpu` need a HIP device (MI355X); all other
tests run on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

import scenes  # noqa: E402

GOLDEN = os.path.join(TESTS, "golden")
GOLDEN_NAMES = ("config1", "wl1", "wl2", "wl3")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def golden_dir(tmp_path_factory):
    """name -> (npz data, scene directory written from it)."""
    base = tmp_path_factory.mktemp("golden_scenes")
    cache = {}

    def get(name):
        if name not in cache:
            z = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
            d = scenes.write_dir(scenes.scene_from_npz(z), str(base / name))
            cache[name] = (z, d)
        return cache[name]

    return get


@pytest.fixture(scope="session")
def workload_dir(tmp_path_factory):
    """(workload, seed) -> scene directory written by the product's generator."""
    base = tmp_path_factory.mktemp("workload_scenes")
    cache = {}

    def get(wl, seed=7):
        key = (wl, seed)
        if key not in cache:
            import gpuprm_amd
            cache[key] = gpuprm_amd.generate_scene(wl, seed, str(base / ("wl%d_s%d" % (wl, seed))))
        return cache[key]

    return get
