import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "trixi.jl_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

# Under pytest-xdist every worker would start one OpenMP thread per CPU in the
# oracle; oversubscribed, its spin-waiting barriers slow the small 2D runs by
# orders of magnitude. One thread per worker (set before libgomp loads).
if os.environ.get("PYTEST_XDIST_WORKER"):
    os.environ.setdefault("OMP_NUM_THREADS", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def kta():
    d = golden("kta.json")
    d.pop("_source")
    return d


def make_oracle(spec):
    from oracle import Oracle
    spec = dict(spec)
    return Oracle(spec.pop("ndims"), spec.pop("polydeg"), spec.pop("level"), spec.pop("cmin"), spec.pop("cmax"),
                  **spec)
