import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _ensure_built():
    """Build the oracle and the engine only if nothing is built yet."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "randomtilingspublic_amd", "librtiling.so")):
        from randomtilingspublic_amd import build as b
        b.build()


_ensure_built()


@pytest.fixture(scope="session")
def orc():
    from tests import oracle_lib
    return oracle_lib


@pytest.fixture(scope="session")
def gpu_ctx():
    import randomtilingspublic_amd as rt
    ctx = rt.Context(0)
    yield ctx
    ctx.close()
