import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU and the built libdiver.so')


@pytest.fixture(scope='session')
def scene64():
    from diver_amd import synthetic as S
    return S.config1_scene()


@pytest.fixture(scope='session')
def cameras():
    from diver_amd import synthetic as S
    return S.hemisphere_cameras()


@pytest.fixture(scope='session')
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch.device('cuda', 0)


def psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float('inf') if mse == 0 else -10.0 * np.log10(mse)
