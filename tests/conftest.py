import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# Moldy's rdf_ptr stand-in must be global before libmoldyforce.so loads
RDF_STUB = os.path.join(ROOT, "tests", "abi_stub", "_build", "librdfstub.so")
if os.path.exists(RDF_STUB):
    import ctypes
    ctypes.CDLL(RDF_STUB, mode=ctypes.RTLD_GLOBAL)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(autouse=True)
def _name_the_case(request):
    """Print the case name before it runs, so a fault names its test."""
    if request.node.get_closest_marker("gpu"):
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
