import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def _ensure_built():
    """Load what is built; build only what is missing."""
    from voxelcraft_menu_amd import _abi
    from voxelcraft_menu_amd import build as B
    if not os.path.exists(_abi.HOST_LIB):
        B.build_host()
    if not os.path.exists(_abi.MARCH_LIB):
        B.build_march()
    # a library that does not match its sources is refused at load (_abi.check_fresh)
    oracle_lib = os.path.join(ROOT, "oracle", "libvxoracle.so")
    if not os.path.exists(oracle_lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"], check=True)


_ensure_built()
