import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Under pytest-xdist every worker loads its own copy of the OpenMP oracle; with
# a full team of threads in each, the per-tick barriers of the many-tick small
# grid tests oversubscribe the cores (a 0.7 s test took 12 minutes with four
# workers).  One thread per worker keeps the parallel run faster than the
# serial one.  (Set before the oracle library, and its OpenMP runtime, loads.)
if os.environ.get("PYTEST_XDIST_WORKER") and "OMP_NUM_THREADS" not in os.environ:
    os.environ["OMP_NUM_THREADS"] = "1"
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def bfa():
    import boltzmannfluids_amd
    boltzmannfluids_amd.load_library()
    return boltzmannfluids_amd
