import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def pytest_collection_modifyitems(config, items):
    # GPU tests are skipped cleanly where no GPU is present (the CPU suite runs with -m "not gpu")
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    have_gpu = os.path.exists("/dev/kfd")
    if have_gpu:
        try:
            import torch
            have_gpu = torch.cuda.is_available()
        except Exception:
            have_gpu = False
    if have_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in gpu_items:
        it.add_marker(skip)
