"""Shared fixtures: the HIP library front end (package dir fmm-fft_amd/) and
the CPU oracle (oracle/, test infrastructure)."""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))


def load_fmmfft_amd():
    """Import the package in fmm-fft_amd/ (a hyphenated directory) as fmmfft_amd."""
    if "fmmfft_amd" in sys.modules:
        return sys.modules["fmmfft_amd"]
    path = os.path.join(REPO, "fmm-fft_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("fmmfft_amd", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fmmfft_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def fmm():
    return load_fmmfft_amd()


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
