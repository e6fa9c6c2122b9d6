"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built HIP
library; everything else runs on the CPU (oracle, host logic, ABI surface)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import cpu

    cpu.lib()
    return cpu


@pytest.fixture(scope="session")
def engine():
    """One HIP context for the whole GPU session.  A missing library or
    device is an error, never a skip: the product has no CPU fallback."""
    from montepath_amd.engine import Engine

    eng = Engine(0)
    info = eng.device_info()
    print(f"\n[engine] {info['device_id']} {info['name']} {info['arch']} "
          f"CUs={info['compute_units']} mem={info['memory_total_mb']}MB", flush=True)
    yield eng
    eng.close()
