"""Shared test fixtures.

Markers: `gpu` tests need an MI355X (they call the HIP library through its C
ABI and compare with the CPU oracle); everything else runs on the CPU.
"""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402  (test infrastructure)

EPS = np.finfo(np.float64).eps


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X); calls the HIP library")


def load_package():
    """Import cuda-cholesky_amd/ (a directory name that is not an identifier)."""
    name = "cuda_cholesky_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "cuda-cholesky_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def cc():
    return load_package()


@pytest.fixture(scope="session")
def gpu(cc):
    """The package, with the GPU checked present (fails loudly otherwise)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    cc.lib()
    return cc


@pytest.fixture(scope="session")
def lapack(gpu):
    h = gpu.LAPACKHandle()
    yield h
    h.close()


@pytest.fixture(scope="session")
def blas(gpu):
    h = gpu.BLASHandle()
    yield h
    h.close()


@pytest.fixture(scope="session")
def orc():
    oracle.lib()
    return oracle
