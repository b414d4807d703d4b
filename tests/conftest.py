import os
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def oracle():
    import kssdtest
    return kssdtest.Oracle()


@pytest.fixture(scope="session")
def hip():
    import kssdtest
    return kssdtest.KssdHip(int(os.environ.get("KSSDHIP_DEVICE", "0")))
