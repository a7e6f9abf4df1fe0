import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SCEN = os.path.join(ROOT, "tichi_sph_amd", "scenarios")
REF_FIXTURES = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def load_scene(name):
    with open(os.path.join(SCEN, name + ".config.json")) as f:
        cb = json.load(f)
    with open(os.path.join(SCEN, name + ".json")) as f:
        sb = json.load(f)
    return cb, sb


@pytest.fixture
def scene():
    return load_scene
