import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "voxel-sim_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")
    config.addinivalue_line("markers", "slow: long-running CPU check")
    lib = os.path.join(ROOT, "oracle", "libvxoracle.so")
    if not os.path.exists(lib) and os.path.exists("/usr/bin/g++"):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=False, capture_output=True)


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def scenes():
    from vxsim import scenes as s
    return s


def _gpu_present() -> bool:
    try:
        from vxsim import lib
        return lib().vx_init() == 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a live device; GPU tests fail loudly without one."""
    from vxsim import lib
    L = lib()
    status = L.vx_init()
    if status != 0:
        pytest.fail("HIP device unavailable: %s" % L.vx_last_error().decode())
    return L


def gpu_present() -> bool:
    return _gpu_present()
