import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def orc_lib():
    import orc
    orc.build()
    return orc.lib()


@pytest.fixture(scope="session")
def ltc_table():
    """(rgba16, rg16) of the project-fitted GGX LTC table, 51 Fresnel layers."""
    from risltc_amd import ltcfit
    rgba, rg, res = ltcfit.table(51)
    return rgba, rg


@pytest.fixture(scope="session")
def small_scene():
    from risltc_amd import synth
    return synth.small_scene(seed=0)


@pytest.fixture(scope="session")
def small_oracle(orc_lib, small_scene, ltc_table):
    import orc
    return orc.Oracle(small_scene, *ltc_table)


def f32_bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
