import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device count only; no HIP context is created)
        return torch.cuda.device_count() > 0
    except Exception:
        return False
