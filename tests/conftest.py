import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    # the CPU oracle is test infrastructure; make sure it is built
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle_mp.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


@pytest.fixture(scope="session")
def golden_phase():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "phase_golden.npz"),
                        allow_pickle=False))
