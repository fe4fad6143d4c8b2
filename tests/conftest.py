import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "needs_ref: compares with the reference build oracle/_ref "
                            "(fails in a GPU test when it is missing, skips in a CPU test)")


def pytest_runtest_setup(item):
    """A GPU parity test that needs the reference build fails when the build
    is missing (a silent skip would hide a whole class of parity checks); a
    CPU test skips, so the suite still runs where flang is absent."""
    if item.get_closest_marker("needs_ref") is None:
        return
    import oracle
    if oracle.reference_available():
        return
    msg = "reference build oracle/_ref/libramses_ref.so is missing (make -C oracle ref)"
    if item.get_closest_marker("gpu") is not None:
        pytest.fail(msg, pytrace=False)
    pytest.skip(msg)


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle.Oracle()


@pytest.fixture(scope="session")
def golden_unsplit():
    import json
    d = np.load(os.path.join(GOLDEN, "unsplit_cases.npz"))
    return json.loads(str(d["meta"])), d


@pytest.fixture(scope="session")
def golden_cmpdt():
    import json
    d = np.load(os.path.join(GOLDEN, "cmpdt_cases.npz"))
    return json.loads(str(d["meta"])), d


@pytest.fixture(scope="session")
def golden_sedov16():
    return np.load(os.path.join(GOLDEN, "sedov16.npz"))


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a real device; fails (does not skip) without one."""
    import torch
    import ramses_amd
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    lib = ramses_amd.load_library()
    torch.zeros(1, device="cuda")
    return lib
