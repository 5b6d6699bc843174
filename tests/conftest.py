import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: large inputs (seconds to minutes)")
    # The oracle is plain C; make sure the checker is built (gcc is present
    # wherever the tests run).
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context for the whole GPU test session."""
    import ascetic_amd as A
    if A.device_count() < 1:
        pytest.fail("no HIP device visible to the gpu tests")
    ctx = A.Context()
    yield ctx
    ctx.close()
