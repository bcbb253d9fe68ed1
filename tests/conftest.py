import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhrt.so on cuda:0)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    """Builds the CPU oracle library once per session (make is a no-op when current)."""
    lib = os.path.join(ROOT, "oracle", "_build", "libhrt_oracle.so")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    except (OSError, subprocess.CalledProcessError):
        if not os.path.exists(lib):
            raise
    yield


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test scheduled but no GPU is visible")
    torch.cuda.init()
    # torch's default stream is the legacy null stream (handle 0), which the
    # engine would read as "use the context stream"; run every torch copy and
    # every engine launch on one explicit stream instead.
    torch.cuda.set_stream(torch.cuda.Stream())
    import heavyrt_amd as H

    ctx = H.Context(0)
    yield ctx
    ctx.close()


def vsgf_path(name: str) -> str:
    return os.path.join(GOLDEN, f"{name}.vsgf")


VSGF_MESHES = ["triangle", "pyramid", "cube", "torus18", "cone_low", "helix_mid"]
