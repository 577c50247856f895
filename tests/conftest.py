"""Shared pytest setup: the `gpu` marker, repo import path and the GPU context fixture."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libctdm.so")


@pytest.fixture(scope="session")
def ctx():
    import torch
    import ctdm_amd

    assert torch.cuda.is_available(), "gpu tests need a GPU"
    c = ctdm_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def orc():
    from oracle import binding

    binding.lib()
    binding.threads(min(16, os.cpu_count() or 1))
    return binding
