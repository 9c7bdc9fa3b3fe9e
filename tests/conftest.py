import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "cpu-free-model_amd")
ORACLE = os.path.join(REPO, "oracle")
for p in (PKG, ORACLE, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

# Load libcfm (and with it the system HIP runtime) before any test module
# imports torch, whose wheel bundles a HIP runtime of its own.
import cfm  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def cfm():
    import cfm as _cfm
    return _cfm


@pytest.fixture(scope="session")
def gpu(cfm):
    if cfm.device_count() < 1:
        pytest.fail("no GPU visible: gpu-marked tests must run on the MI355X box")
    return cfm


@pytest.fixture(autouse=True)
def _announce_gpu_case(request):
    # name the case (and flush) before any GPU work, so a run that dies mid-way
    # still says which case it was on
    if request.node.get_closest_marker("gpu") is not None:
        sys.stderr.write(f"\n[gpu case] {request.node.nodeid}\n")
        sys.stderr.flush()
    yield
