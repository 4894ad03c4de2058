import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this container")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def small_survey():
    """150 forests of the 10x10 deg recipe at c0 density (seed 7)."""
    from lya2pcf_amd import synth
    return synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 150, seed=7))


@pytest.fixture(scope="session")
def small_set(small_survey):
    from lya2pcf_amd import forests
    return forests.ForestSet.from_survey(small_survey)


@pytest.fixture(scope="session")
def small_pairs(small_set):
    from lya2pcf_amd import forests, xi
    am = forests.angmax_from(small_set, 200.0)
    return am, xi.build_pairs(small_set, am)
