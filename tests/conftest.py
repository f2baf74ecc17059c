import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "nest-simulator_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the device path")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available():
    try:
        import nestgpu
        return nestgpu.device_count() > 0
    except ImportError:
        return False
