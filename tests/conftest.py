import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running GPU case")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle()


@pytest.fixture(scope="session")
def pg():
    import perft_gpu_amd
    if not os.path.exists(perft_gpu_amd.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "perft_gpu_amd"), "-j8"], check=True)
    perft_gpu_amd.lib()
    return perft_gpu_amd


@pytest.fixture(scope="session")
def dev(pg):
    """Context that times every leaf launch: pg_perft runs the host-driven level loop."""
    d = pg.Device(0, arena_bytes=24 << 30, timing=True)
    yield d
    d.close()


@pytest.fixture(scope="session")
def gdev(pg):
    """Context without per-launch timing: pg_perft runs as one graph launch."""
    d = pg.Device(0, arena_bytes=16 << 30)
    yield d
    d.close()
