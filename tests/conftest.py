import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "huntian.jl_amd"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and libhuntian_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def websky():
    from huntian_hip import WEBSKY_TABLE, read_homeltab

    return read_homeltab(WEBSKY_TABLE)


@pytest.fixture(scope="session")
def fortran_table():
    from huntian_hip import read_homeltab

    return read_homeltab(os.path.join(REPO, "tests", "golden", "HomelTab_fortran.dat"))
