import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def nerf():
    return importlib.import_module("nerf-tf2_amd")


@pytest.fixture(scope="session")
def oracle():
    return importlib.import_module("oracle.restatement")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
