import os
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    # kernel shapes stay the handles' defaults: no tuning cache is read at
    # index open, and cache writes go to a file of this session only
    os.environ["KSSDHIP_NO_TUNE_CACHE"] = "1"
    import tempfile
    os.environ["KSSDHIP_TUNE_CACHE"] = str(Path(tempfile.mkdtemp(prefix="kssdhip_tune_")) / "tuning.json")


@pytest.fixture(scope="session")
def oracle():
    import kssdtest
    return kssdtest.Oracle()


@pytest.fixture(scope="session")
def hip():
    import kssdtest
    return kssdtest.KssdHip(int(os.environ.get("KSSDHIP_DEVICE", "0")))
