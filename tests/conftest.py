import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def oracle_link(cfg, detector=None):
    """oracle.drop.Link with the parameters of a hermespy_amd LinkConfig."""

    from oracle import drop, ofdm

    frame = None
    if cfg.frame_mask is not None:
        frame = ofdm.Frame(cfg.num_subcarriers, np.full(cfg.words, cfg.prefix, dtype=int), cfg.frame_mask.copy(),
                           cfg.dc_suppression)
    return drop.Link(cfg.num_rx, cfg.num_tx, cfg.num_subcarriers, cfg.words, cfg.prefix_ratio, cfg.modulation_order,
                     cfg.bandwidth, cfg.delays, cfg.power_profile, cfg.rice_factors, cfg.doppler, cfg.num_sinusoids,
                     cfg.corr_rx, cfg.corr_tx, detector or cfg.detector, cfg.dc_suppression, cfg.gain,
                     cfg.device_power, frame, cfg.pilots, cfg.los_doppler)


@pytest.fixture(scope="session")
def native_lib():
    """The built library (built here if absent, so CPU tests see the current sources)."""

    from hermespy_amd import _lib

    if not _lib.LIB_PATH.exists():
        from hermespy_amd.build import build

        build()
    return _lib.load()


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
