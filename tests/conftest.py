import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def small_scene():
    from probabilistic_mapping_amd import synth
    return synth.small_scene(0)


@pytest.fixture(scope="session")
def trajectory_scene():
    from probabilistic_mapping_amd import synth
    return synth.trajectory_scene(0)
