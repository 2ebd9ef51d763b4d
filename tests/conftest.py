import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "krakenn.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    from tests import _records

    _records.set_writer(config)


@pytest.fixture(scope="session")
def po():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def kmi():
    import kraken_mi355x

    kraken_mi355x.load()
    return kraken_mi355x


@pytest.fixture(scope="session")
def backend(kmi):
    return kmi.MI355XBackend(0)
