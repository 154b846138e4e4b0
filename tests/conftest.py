import os
import pathlib
import sys

# Under pytest-xdist every worker loads the OpenMP oracle library; with a full
# thread team each and spin-waiting between its many short parallel regions
# the workers starve one another.  Split the cores and wait passively
# (before any test loads the library).
if "PYTEST_XDIST_WORKER" in os.environ:
    _workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1"))
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // max(1, _workers)))
    os.environ["OMP_WAIT_POLICY"] = "PASSIVE"

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running validation")
