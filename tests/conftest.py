import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libtc3.so")


@pytest.fixture(scope="session")
def inputs_dir(tmp_path_factory):
    """Generated Morse inputs of the workload families, written once per session."""
    from tdep_amd import inputs
    base = tmp_path_factory.mktemp("inputs")
    cache = {}

    def get(name):
        if name not in cache:
            d = str(base / name)
            inputs.make_config(name, d)
            cache[name] = d
        return cache[name]
    return get


@pytest.fixture(scope="session")
def oracle_setup(inputs_dir):
    from oracle import run as R
    cache = {}

    def get(name, grid):
        key = (name, tuple(grid))
        if key not in cache:
            cache[key] = R.setup(inputs_dir(name), grid)
        return cache[key]
    return get
