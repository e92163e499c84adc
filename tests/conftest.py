import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "sph-exa-nnet_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import Reference, reference_available
    if not reference_available():
        pytest.skip("reference build oracle/_ref not present (built only where the reference tree exists)")
    return Reference()
