"""Shared test setup: the `gpu` marker, import paths and library handles."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "hvm-64_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def load_snapshots():
    with open(os.path.join(GOLDEN, "run_snapshots.json")) as f:
        snaps = json.load(f)
    for v in snaps.values():
        with open(os.path.join(GOLDEN, v["program"])) as f:
            v["text"] = f.read()
    return snaps


@pytest.fixture(scope="session")
def snapshots():
    return load_snapshots()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def gpu_heap():
    """One engine handle (16 GiB heap) shared by the GPU tests."""
    import hvm64
    heap = hvm64.Heap(16 << 30)
    heap.set_timeout(120)
    yield heap
    heap.close()
