import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP path through the C ABI")


@pytest.fixture(scope="session")
def oracle_esm():
    import oracle_lib

    lib = oracle_lib.load("esm")
    if lib is None:
        pytest.fail("oracle/liboracle_esm.so is missing: run `make -C oracle`")
    return lib


@pytest.fixture(scope="session")
def oracle_libm():
    import oracle_lib

    lib = oracle_lib.load("libm")
    if lib is None:
        pytest.fail("oracle/liboracle_libm.so is missing: run `make -C oracle`")
    return lib


def announce(msg):
    """Names the GPU case before it runs (so a failing run shows what it was doing)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()
