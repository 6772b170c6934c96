import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "build", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "pmx_oracle.cpp")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       stdout=subprocess.DEVNULL)


_ensure_oracle()


@pytest.fixture(scope="session")
def manifest():
    with open(os.path.join(HERE, "golden", "manifest.json")) as fh:
        return json.load(fh)
